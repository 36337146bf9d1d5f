// Cell-centred field dumps as text: the reference's output (:4834-4883)
// and output_tecplot (:4885-4938) for one subdomain.
//
// The planes are downloaded as doubles (either numeric form), the columns
// are evaluated with the reference's expressions in its order of
// operations, and the text of the k-planes is formatted by several host
// threads and written in order.  phi is the potential of the last Poisson
// correction (0 unless poisson_every enables renew_phi3, which the
// reference compiles out).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <thread>
#include <vector>

#include "../../include/plasmasim.h"

namespace {

struct Planes {
  int SX, SY, SZ;
  std::vector<double> f[PIC_NFIELDS + 1];  // + PIC_PHI
  double at(int w, int k, int l, int m) const {
    return f[w][((size_t)k * SY + l) * SZ + m];
  }
};

void put(std::string &s, double v, bool first) {
  char b[32];
  const int n = snprintf(b, sizeof b, first ? "%.3E" : " %.3E", v);
  s.append(b, (size_t)n);
}

// one line of output() (:4846-4876) for cell (k,l,m)
void line_columns(const Planes &p, const pic_config &c, int k, int l, int m,
                  std::string &s) {
  const double dx = c.dx;
  const int Nx = c.nx, Ny = c.ny, Nz = c.nz;
  double v[23];
  v[0] = (k - 2 - Nx / 2) * dx;
  v[1] = (l - 2 - Ny / 2) * dx;
  v[2] = (m - (Nz + 4) / 2) * dx;
  v[3] = p.at(PIC_NI, k, l, m);
  v[4] = p.at(PIC_NE, k, l, m);
  v[5] = p.at(PIC_NI, k, l, m) - p.at(PIC_NE, k, l, m);
  v[6] = (p.at(PIC_EX, k + 1, l, m) + p.at(PIC_EX, k, l, m)) / 2.;
  v[7] = (p.at(PIC_EY, k, l + 1, m) + p.at(PIC_EY, k, l, m)) / 2.;
  v[8] = (p.at(PIC_EZ, k, l, m + 1) + p.at(PIC_EZ, k, l, m)) / 2.;
  const double bx = (p.at(PIC_BX, k, l, m) + p.at(PIC_BX, k, l + 1, m) +
                     p.at(PIC_BX, k, l, m + 1) + p.at(PIC_BX, k, l + 1, m + 1)) / 4.;
  const double by = (p.at(PIC_BY, k, l, m) + p.at(PIC_BY, k + 1, l, m) +
                     p.at(PIC_BY, k, l, m + 1) + p.at(PIC_BY, k + 1, l, m + 1)) / 4.;
  const double bz = (p.at(PIC_BZ, k, l, m) + p.at(PIC_BZ, k + 1, l, m) +
                     p.at(PIC_BZ, k, l + 1, m) + p.at(PIC_BZ, k + 1, l + 1, m)) / 4.;
  const double b0x = (p.at(PIC_B0X, k, l, m) + p.at(PIC_B0X, k, l + 1, m) +
                      p.at(PIC_B0X, k, l, m + 1) + p.at(PIC_B0X, k, l + 1, m + 1)) / 4.;
  const double b0y = (p.at(PIC_B0Y, k, l, m) + p.at(PIC_B0Y, k + 1, l, m) +
                      p.at(PIC_B0Y, k, l, m + 1) + p.at(PIC_B0Y, k + 1, l, m + 1)) / 4.;
  const double b0z = (p.at(PIC_B0Z, k, l, m) + p.at(PIC_B0Z, k + 1, l, m) +
                      p.at(PIC_B0Z, k, l + 1, m) + p.at(PIC_B0Z, k + 1, l + 1, m)) / 4.;
  v[9] = bx;
  v[10] = by;
  v[11] = bz;
  for (int j = 0; j < 6; ++j) v[12 + j] = p.at(PIC_JIX + j, k, l, m);
  v[18] = p.at(PIC_PHI, k, l, m);
  v[19] = sqrt(bx * bx + by * by + bz * bz);
  v[20] = bx - b0x;
  v[21] = by - b0y;
  v[22] = bz - b0z;
  for (int j = 0; j < 23; ++j) put(s, v[j], j == 0);
  s.push_back('\n');
}

// one line of output_tecplot() (:4896-4930): J averaged to the cell centre
// and the magnitude of J_i - J_e added
void line_tecplot(const Planes &p, const pic_config &c, int k, int l, int m,
                  std::string &s) {
  const double dx = c.dx;
  const int Nx = c.nx, Ny = c.ny, Nz = c.nz;
  double v[24];
  v[0] = (k - 2 - Nx / 2) * dx;
  v[1] = (l - 2 - Ny / 2) * dx;
  v[2] = (m - (Nz + 4) / 2) * dx;
  v[3] = p.at(PIC_NI, k, l, m);
  v[4] = p.at(PIC_NE, k, l, m);
  v[5] = p.at(PIC_NI, k, l, m) - p.at(PIC_NE, k, l, m);
  v[6] = (p.at(PIC_EX, k + 1, l, m) + p.at(PIC_EX, k, l, m)) / 2.;
  v[7] = (p.at(PIC_EY, k, l + 1, m) + p.at(PIC_EY, k, l, m)) / 2.;
  v[8] = (p.at(PIC_EZ, k, l, m + 1) + p.at(PIC_EZ, k, l, m)) / 2.;
  const double bx = (p.at(PIC_BX, k, l, m) + p.at(PIC_BX, k, l + 1, m) +
                     p.at(PIC_BX, k, l, m + 1) + p.at(PIC_BX, k, l + 1, m + 1)) / 4.;
  const double by = (p.at(PIC_BY, k, l, m) + p.at(PIC_BY, k + 1, l, m) +
                     p.at(PIC_BY, k, l, m + 1) + p.at(PIC_BY, k + 1, l, m + 1)) / 4.;
  const double bz = (p.at(PIC_BZ, k, l, m) + p.at(PIC_BZ, k + 1, l, m) +
                     p.at(PIC_BZ, k, l + 1, m) + p.at(PIC_BZ, k + 1, l + 1, m)) / 4.;
  const double b0x = (p.at(PIC_B0X, k, l, m) + p.at(PIC_B0X, k, l + 1, m) +
                      p.at(PIC_B0X, k, l, m + 1) + p.at(PIC_B0X, k, l + 1, m + 1)) / 4.;
  const double b0y = (p.at(PIC_B0Y, k, l, m) + p.at(PIC_B0Y, k + 1, l, m) +
                      p.at(PIC_B0Y, k, l, m + 1) + p.at(PIC_B0Y, k + 1, l, m + 1)) / 4.;
  const double b0z = (p.at(PIC_B0Z, k, l, m) + p.at(PIC_B0Z, k + 1, l, m) +
                      p.at(PIC_B0Z, k, l + 1, m) + p.at(PIC_B0Z, k + 1, l + 1, m)) / 4.;
  v[9] = bx;
  v[10] = by;
  v[11] = bz;
  // face-centred J components averaged along their own axis
  const int dk[3] = {1, 0, 0}, dl[3] = {0, 1, 0}, dm[3] = {0, 0, 1};
  double ji[3], je[3];
  for (int a = 0; a < 3; ++a) {
    ji[a] = (p.at(PIC_JIX + a, k + dk[a], l + dl[a], m + dm[a]) +
             p.at(PIC_JIX + a, k, l, m)) / 2.;
    je[a] = (p.at(PIC_JEX + a, k + dk[a], l + dl[a], m + dm[a]) +
             p.at(PIC_JEX + a, k, l, m)) / 2.;
    v[12 + a] = ji[a];
    v[15 + a] = je[a];
  }
  v[18] = sqrt((ji[0] - je[0]) * (ji[0] - je[0]) + (ji[1] - je[1]) * (ji[1] - je[1]) +
               (ji[2] - je[2]) * (ji[2] - je[2]));
  v[19] = p.at(PIC_PHI, k, l, m);
  v[20] = sqrt(bx * bx + by * by + bz * bz);
  v[21] = bx - b0x;
  v[22] = by - b0y;
  v[23] = bz - b0z;
  for (int j = 0; j < 24; ++j) put(s, v[j], j == 0);
  s.push_back('\n');
}

}  // namespace

extern "C" int pic_output_fields(pic_ctx *ctx, const char *path, int format,
                                 int64_t step) {
  if (!ctx || !path || (format != PIC_OUTPUT_COLUMNS && format != PIC_OUTPUT_TECPLOT))
    return PIC_E_ARG;
  pic_config c;
  int rc = pic_get_config(ctx, &c);
  if (rc) return rc;
  Planes p;
  p.SX = c.nx + 4;
  p.SY = c.ny + 4;
  p.SZ = c.nz + 4;
  const size_t nc = (size_t)p.SX * p.SY * p.SZ;
  try {
    for (int w = 0; w <= PIC_PHI; ++w) p.f[w].resize(nc);
  } catch (...) {
    return PIC_E_NOMEM;
  }
  for (int w = 0; w <= PIC_PHI; ++w)
    if ((rc = pic_download_field(ctx, w, p.f[w].data()))) return rc;

  const bool tec = format == PIC_OUTPUT_TECPLOT;
  const int k1 = tec ? c.nx + 3 : c.nx + 2, l1 = tec ? c.ny + 3 : c.ny + 2;
  const int nk = k1 - 2;
  std::vector<std::string> text((size_t)nk);
  unsigned nthreads = std::thread::hardware_concurrency();
  if (nthreads == 0) nthreads = 4;
  if (nthreads > 32) nthreads = 32;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthreads; ++t)
    pool.emplace_back([&, t]() {
      for (int k = 2 + (int)t; k < k1; k += (int)nthreads) {
        std::string &s = text[(size_t)(k - 2)];
        for (int l = 2; l < l1; ++l) {
          for (int m = 2; m < c.nz + 2; ++m) {
            if (tec) line_tecplot(p, c, k, l, m, s);
            else line_columns(p, c, k, l, m, s);
          }
          if (tec) s.push_back('\n');
        }
      }
    });
  for (auto &th : pool) th.join();

  FILE *f = fopen(path, "w");
  if (!f) return PIC_E_IO;
  bool ok = true;
  if (tec) {
    ok = fprintf(f,
                 "VARIABLES = \"X[m]\" \"Y[m]\" \"Z[m]\" \"ni[/m3]\" \"ne[/m3]\" "
                 "\"dens[/m3]\" \"Ex[V/m]\" \"Ey[V/m]\" \"Ez[V/m]\" \"Bx[T]\" "
                 "\"By[T]\" \"Bz[T]\" \"jix[A/m2]\" \"jiy[A/m2]\" \"jiz[A/m2]\" "
                 "\"jex[A/m2]\" \"jey[A/m2]\" \"jez[A/m2]\" \"j[A/m2]\" "
                 "\"phi[V]\" \"B[T]\" \"Bpx[T]\" \"Bpy[T]\" \"Bpz[T]\"\n") > 0 &&
         fprintf(f,
                 "ZONE T=\"STP:%lld\", STRANDID=1, SOLUTIONTIME=%lld, I=%d, J=%d, "
                 "K=%d\n",
                 (long long)step, (long long)step, c.nz, c.ny + 1, c.nx + 1) > 0;
  }
  for (int k = 0; k < nk && ok; ++k)
    ok = fwrite(text[(size_t)k].data(), 1, text[(size_t)k].size(), f) ==
         text[(size_t)k].size();
  if (fclose(f) != 0) ok = false;
  return ok ? 0 : PIC_E_IO;
}
