// sweep_c_abi.cpp — a C++ caller of libqss through include/qss.h only, the
// way the reference's host code (Mesh::MoveAllSplitLayers, step.cpp:105)
// would call it: create a context, sweep, read the Ex history, destroy.
//
// usage: sweep_c_abi <out.bin>
// writes nslices residuals followed by the nslices x nz x ny Ex history
// (float64, native byte order) of a 32x32, 20-layer sweep.
#include <cstdio>
#include <vector>

#include "qss.h"

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <out.bin>\n", argv[0]);
    return 2;
  }
  qss_params p = {};
  p.ny = 32;
  p.nz = 32;
  p.ly = 6.0;
  p.lz = 6.0;
  p.hx = 0.05;
  p.nx_split = 1;
  p.nslices = 20;
  p.n_iter = 2;
  p.dens = 1.0;
  p.ppc_y = 2;
  p.ppc_z = 2;
  p.jitter = 1e-3;
  p.seed = 12345;
  p.shape = QSS_SHAPE_CIC;
  p.fix_subcycle = 0;
  p.precision = QSS_PREC_FP64;
  p.nbunch = 1;
  p.bunch[0].nb = 2.0;
  p.bunch[0].sigma_r = 0.5;
  p.bunch[0].sigma_xi = 0.3;
  p.bunch[0].xi0 = 0.5;
  p.bunch[0].polarity = -1.0;
  p.bunch[0].vx = 1.0;
  p.nwake = 1;

  qss_ctx *c = nullptr;
  if (qss_create(&p, 0, 1u << QSS_EX, &c) != QSS_OK) {
    std::fprintf(stderr, "qss_create: %s\n", qss_last_error());
    return 1;
  }
  std::vector<double> res((size_t)p.nslices * p.nx_split);
  int st = qss_sweep(c, res.data());
  if (st != QSS_OK) {
    std::fprintf(stderr, "qss_sweep: %d %s\n", st, qss_last_error());
    qss_destroy(c);
    return 1;
  }
  const size_t plane = (size_t)p.ny * p.nz;
  std::vector<double> ex(plane * p.nslices);
  for (int s = 0; s < p.nslices && st == QSS_OK; ++s)
    st = qss_get_history(c, QSS_EX, s, ex.data() + plane * s);
  qss_destroy(c);
  if (st != QSS_OK) {
    std::fprintf(stderr, "qss_get_history: %s\n", qss_last_error());
    return 1;
  }
  std::FILE *f = std::fopen(argv[1], "wb");
  if (!f) {
    std::perror(argv[1]);
    return 1;
  }
  const bool ok = std::fwrite(res.data(), sizeof(double), res.size(), f) == res.size() &&
                  std::fwrite(ex.data(), sizeof(double), ex.size(), f) == ex.size();
  std::fclose(f);
  if (!ok) return 1;
  std::printf("swept %d layers, last residual %.3e\n", p.nslices, res.back());
  return 0;
}
