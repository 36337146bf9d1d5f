"""The oracle's output() / output_tecplot() dumps against closed forms: with
fields linear in the cell indices every cell average is known exactly."""
import io

import numpy as np
import pytest

import plasmasim_amd as ps
from oracle_lib import Oracle


def linear_state():
    c = ps.preset(0)
    o = Oracle(c)
    S = o.shape
    k, l, m = np.meshgrid(*[np.arange(n, dtype=float) for n in S], indexing="ij")
    f = {"ex": k, "ey": 2 * l, "ez": 3 * m, "bx": l + m, "by": k + m, "bz": k + l,
         "b0x": 0.5 * (l + m), "b0y": 0.5 * (k + m), "b0z": 0.5 * (k + l),
         "jix": k + 0.25, "jiy": l + 0.5, "jiz": m + 0.75, "jex": 0.1 * k,
         "jey": 0.2 * l, "jez": 0.3 * m, "ni": 1 + k, "ne": 0.5 + 0 * k}
    for n, a in f.items():
        o.upload_field(n, a)
    return c, o


def expected(c, k, l, m, tec):
    dx = c.dx
    x = [(k - 2 - c.nx // 2) * dx, (l - 2 - c.ny // 2) * dx, (m - (c.nz + 4) // 2) * dx]
    ni, ne = 1 + k, 0.5
    e = [k + 0.5, 2 * l + 1, 3 * m + 1.5]
    b = [l + m + 1, k + m + 1, k + l + 1]
    if tec:
        ji = [k + 0.75, l + 1.0, m + 1.25]
        je = [0.1 * (k + 0.5), 0.2 * (l + 0.5), 0.3 * (m + 0.5)]
        j = [*ji, *je, np.sqrt(sum((a - bb) ** 2 for a, bb in zip(ji, je)))]
        tail = [0.0]
    else:
        j = [k + 0.25, l + 0.5, m + 0.75, 0.1 * k, 0.2 * l, 0.3 * m]
        tail = [0.0]
    bmag = np.sqrt(sum(v * v for v in b))
    return [*x, ni, ne, ni - ne, *e, *b, *j, *tail, bmag, *[0.5 * v for v in b]]


@pytest.mark.parametrize("tec", [False, True])
def test_dump_columns_follow_closed_forms(tmp_path, tec):
    c, o = linear_state()
    path = tmp_path / "f.dat"
    o.output_fields(str(path), tecplot=tec, step=42)
    lines = path.read_text().split("\n")
    if tec:
        assert lines[0].startswith("VARIABLES = ") and lines[0].count('"') == 48
        assert lines[1] == (f'ZONE T="STP:42", STRANDID=1, SOLUTIONTIME=42, '
                            f"I={c.nz}, J={c.ny + 1}, K={c.nx + 1}")
        body = lines[2:]
        cells = [(k, l, m) for k in range(2, c.nx + 3) for l in range(2, c.ny + 3)
                 for m in range(2, c.nz + 2)]
        # an empty line after each (k, l) column of cells
        blanks = [i for i, s in enumerate(body) if s == ""]
        assert len(blanks) == (c.nx + 1) * (c.ny + 1) + 1  # + the final newline
        rows = [s for s in body if s]
    else:
        cells = [(k, l, m) for k in range(2, c.nx + 2) for l in range(2, c.ny + 2)
                 for m in range(2, c.nz + 2)]
        rows = [s for s in lines if s]
    assert len(rows) == len(cells)
    ncol = 24 if tec else 23
    for s in rows[:3]:
        assert all(len(t) in (9, 10) and "E" in t for t in s.split(" "))
    got = np.loadtxt(io.StringIO("\n".join(rows)))
    assert got.shape == (len(cells), ncol)
    want = np.array([expected(c, *kl, tec) for kl in cells])
    # printed with 4 significant digits
    assert np.allclose(got, want, rtol=6e-4, atol=1e-12)
