"""Checkpoint format on the CPU: the oracle's writer against the layout and
hash that include/plasmasim.h specifies (re-implemented here from the
specification), save/resume round trips, and rejected files."""
import os
import struct

import numpy as np
import pytest

import plasmasim_amd as ps
from oracle_lib import Oracle

FIELDS = ["ex", "ey", "ez", "bx", "by", "bz", "b0x", "b0y", "b0z", "jix", "jiy",
          "jiz", "jex", "jey", "jez", "ni", "ne"]
HDR = struct.Struct("<8sIIIi3i3iqiiII2q2qQQ136s")
MASK = (1 << 64) - 1


def spec_hash(payload):
    """The payload hash as specified next to pic_checkpoint_header."""
    pad = (-len(payload)) % 8
    words = np.frombuffer(payload + b"\0" * pad, dtype="<u8")
    h = [(0xcbf29ce484222325 + j) & MASK for j in range(4)]
    for i, w in enumerate(words.tolist()):
        j = i & 3
        h[j] = ((h[j] ^ w) * 0x100000001b3) & MASK
    return ((((h[0] * 31 + h[1]) * 31 + h[2]) * 31) + h[3]) & MASK


def cfg0():
    c = ps.preset(0)
    c.inflow_every = 3
    c.sort_every = 4
    return c


def state(o):
    f = {n: o.download_field(n) for n in FIELDS}
    p = [o.download_particles(sp) for sp in (0, 1)]
    return f, p


def assert_same(a, b):
    for n in FIELDS:
        assert np.array_equal(a[0][n], b[0][n], equal_nan=True), n
    for sp in (0, 1):
        assert a[1][sp].shape == b[1][sp].shape
        assert np.array_equal(a[1][sp], b[1][sp], equal_nan=True), sp


@pytest.fixture
def saved(tmp_path):
    c = cfg0()
    o = Oracle(c)
    o.load_particles()
    o.begin()
    o.step(5)
    path = str(tmp_path / "o.ckpt")
    assert o.checkpoint_save(path) == 0
    return c, o, path


def test_layout_and_hash_follow_the_specification(saved):
    c, o, path = saved
    raw = open(path, "rb").read()
    assert HDR.size == 256
    h = HDR.unpack(raw[:256])
    (magic, version, hbytes, cbytes, prec, nx, ny, nz, tx, ty, tz, step, c_add,
     c_sort, event, ntiles, n0, n1, se0, se1, pbytes, csum, _) = h
    assert magic == b"PLASCKPT" and version == 1 and hbytes == 256
    assert cbytes == ps.lib().pic_sizeof_config()
    assert prec == ps.PIC_FP64 and (nx, ny, nz) == (c.nx, c.ny, c.nz)
    assert step == 5 and ntiles == 0
    assert (n0, n1) == (o.particle_count(0), o.particle_count(1))
    payload = raw[256:]
    assert pbytes == len(payload)
    nc = (c.nx + 4) * (c.ny + 4) * (c.nz + 4)
    assert len(payload) == cbytes + 17 * nc * 8 + 6 * 8 * (n0 + n1)
    assert spec_hash(payload) == csum
    # planes in enum order, then six arrays per species
    off = cbytes
    ex = np.frombuffer(payload, "<f8", nc, off).reshape(o.shape)
    assert np.array_equal(ex, o.download_field("ex"))
    off += 17 * nc * 8
    p = o.download_particles(0)
    z = np.frombuffer(payload, "<f8", n0, off + 2 * n0 * 8)
    assert np.array_equal(z, p[:, 2], equal_nan=True)


def test_resume_continues_bit_identically(saved):
    c, o, path = saved
    r = Oracle(c)
    assert r.checkpoint_load(path) == 0
    assert_same(state(o), state(r))
    for x in (o, r):
        x.step(7)  # crosses inflow and sort cadences
    assert_same(state(o), state(r))
    assert o.diagnose()["count"].tolist() == r.diagnose()["count"].tolist()


def test_corrupt_truncated_or_foreign_files_are_rejected(saved, tmp_path):
    c, o, path = saved
    raw = bytearray(open(path, "rb").read())
    bad = tmp_path / "bad.ckpt"
    flipped = bytearray(raw)
    flipped[-9] ^= 0x40
    bad.write_bytes(bytes(flipped))
    assert Oracle(c).checkpoint_load(str(bad)) == -7
    bad.write_bytes(bytes(raw[:len(raw) // 2]))
    assert Oracle(c).checkpoint_load(str(bad)) in (-6, -7)
    bad.write_bytes(b"NOTACKPT" + bytes(raw[8:]))
    assert Oracle(c).checkpoint_load(str(bad)) == -7
    other = cfg0()
    other.seed += 1
    assert Oracle(other).checkpoint_load(path) == -7
    assert Oracle(c).checkpoint_load(str(tmp_path / "missing.ckpt")) == -6


def test_capacity_and_deposit_mode_may_differ(saved):
    c, o, path = saved
    d = cfg0()
    d.capacity = 3 * d.nx * d.ny * d.nz * d.ppc
    d.deposit = ps.PIC_DEPOSIT_DETERMINISTIC
    r = Oracle(d)
    assert r.checkpoint_load(path) == 0
    assert_same(state(o), state(r))
    os.remove(path)
