// save.cpp — qss_save / qss_load: the reference's HDF5 output and restart
// files (Mesh::Save, save.cpp:178-835; Mesh::Load, load.cpp:285-1080) written
// from and read into a context, one process (PE 0) holding the whole grid.
//
// vsNNN_3d_fields.h5
//   meshinfo  I32[16]  the 16 ints that follow &l_Mx in Mesh (mesh.h:16-19):
//                      Mx dMx sizeX offsetX My dMy sizeY offsetY Mz dMz sizeZ
//                      offsetZ, then the two 32-bit halves of sizeXY and of
//                      sizeXZ (longs), with ghost widths dM = 2 (mesh.h:232)
//   phase     F64[1]
//   X, Y, Z   F64 axes: (i + shift) hx, j hy - Ly/2, k hz - Lz/2 (save.cpp:540-575)
//   ex ey ez bx by bz jx jy jz dens Dens0 .. Dens<nsorts-1>
//             F64[Mz][My][Mx], element (k, j, i) at i + Mx (j + My k)
// The 3-D cell i of the reference holds layer iLayer = i after the sweep
// (movepart_splitlayer.cpp:467-472), which is history slice Mx-1-i here.
// "dens" and "Dens<s>" are what Mesh::Density(ALLSORTS) deposits from the
// particles of the 3-D mesh (mesh.cpp:712-760, Cell3Dm::AddDensity,
// cell3d.cpp:6-43): the beam.  With an evolving driver they are deposited
// here from its particles (charge state 1; deposits into ghost cells are not
// folded back); with the frozen driver, "dens" is the beam charge density
// of each layer (the rho_b history) and Dens1 its magnitude.  Sort 0, the
// plasma electrons, lives in the layers, not in the 3-D mesh: Dens0 is zero.
//
// vsNNN_3d_particles.h5
//   Ions_of_sort_%02d_at_PE_0  SavedIon records (save.cpp:256-281) of the
//                      driver particles of bunch s-1 in interior cells, in
//                      cell order (k, j, i) as Mesh::Save visits them: cell =
//                      the ghosted cell number GetN(i,j,k) (mesh.h:150), x =
//                      (i + fx) hx, y = (j + fy) hy, z = (k + fz) hz,
//                      momenta, weight, q2m of the bunch, i_z = 1, sort = s
//   Nparticles_at_PE_0  I64[nsorts]  particles written per sort
// Sorts: 0 the plasma (no particles in the 3-D mesh), 1 + b bunch b of an
// evolving driver; with the frozen driver there is one beam sort.

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "h5file.h"
#include "qss_internal.h"

namespace {

constexpr int kGhost = 2;   // l_dMx = l_dMy = l_dMz (mesh.h:232)
const char *kFieldName[10] = {"ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "dens"};
const int kFieldId[9] = {QSS_EX, QSS_EY, QSS_EZ, QSS_BX, QSS_BY, QSS_BZ, QSS_JX, QSS_JY, QSS_JZ};

int fail(const std::string &m, int code = QSS_E_ARG) {
  qss_set_error(m);
  return code;
}

std::string file_name(const char *dir, int isave, const char *what) {
  char b[64];
  snprintf(b, sizeof b, "vs%3.3d_3d_%s.h5", isave, what);
  std::string d = dir && *dir ? dir : ".";
  return d + "/" + b;
}

struct Grid {
  long L, ny, nz;
  long sx() const { return L + 2 * kGhost; }
  long sy() const { return ny + 2 * kGhost; }
  long sz() const { return nz + 2 * kGhost; }
  long cell(long i, long j, long k) const {
    return i + kGhost + sx() * (j + kGhost + sy() * (k + kGhost));
  }
};

// history of field f for the selected wake, slices [0, L), on the host
int fetch_history(qss_ctx *c, int f, std::vector<double> &h) {
  const size_t n = (size_t)c->p.nslices * c->ncell;
  h.resize(n);
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  HIPCHECK(hipMemcpy(h.data(), c->hist[f] + (size_t)c->sel * n, n * sizeof(double),
                     hipMemcpyDeviceToHost));
  return QSS_OK;
}

// one [Mz][My][Mx] field dataset from slice-major planes (plane s = layer L-1-s)
bool put_field(h5::Writer &w, int set, const Grid &g, const std::vector<double> &hist) {
  std::vector<double> row((size_t)g.ny * g.L);
  const size_t nc = (size_t)g.ny * g.nz;
  for (long k = 0; k < g.nz; ++k) {
    for (long j = 0; j < g.ny; ++j)
      for (long i = 0; i < g.L; ++i)
        row[(size_t)j * g.L + i] = hist[(size_t)(g.L - 1 - i) * nc + j + (size_t)g.ny * k];
    if (!w.put(set, (uint64_t)k * row.size() * sizeof(double), row.data(),
               row.size() * sizeof(double)))
      return false;
  }
  return true;
}

// the same from a [k][j][i] array (host deposits)
bool put_cells(h5::Writer &w, int set, const std::vector<double> &v) {
  return w.put(set, 0, v.data(), v.size() * sizeof(double));
}

struct DriverHost {
  long n = 0;
  std::vector<int32_t> cell, bunch;
  std::vector<double> frac, mom, w;
};

}  // namespace

extern "C" int qss_save(qss_ctx *c, const char *dir, int isave, double phase, int shift) {
  if (!c) return fail("save: null context");
  if (isave < 0 || isave > 999) return fail("save: isave must be in [0, 999] (vs%3.3d names)");
  const Grid g{c->p.nslices, c->ny, c->nz};
  std::string missing;
  for (int f = 0; f < 9; ++f)
    if (!c->hist[kFieldId[f]]) missing += std::string(" ") + kFieldName[f];
  const bool evolving = c->drv.on;
  if (!evolving && !c->hist[QSS_RHOBEAM]) missing += " rhobeam (dens of the frozen driver)";
  if (!missing.empty())
    return fail("save: the context keeps no history of:" + missing +
                " (set those bits of history_mask at create)");

  // driver particles in interior cells, cell order (k, j, i), stable
  DriverHost d;
  const int nb = c->p.nbunch;
  const int nsorts = evolving ? 1 + nb : 2;
  std::vector<long> order;
  if (evolving) {
    int64_t n = c->drv.np;
    d.n = n;
    d.cell.resize(3 * n); d.bunch.resize(n); d.frac.resize(3 * n); d.mom.resize(3 * n); d.w.resize(n);
    CHECK(qss_driver_get(c, d.cell.data(), d.frac.data(), d.mom.data(), d.w.data(), d.bunch.data()));
    std::vector<long> key(n);
    for (long q = 0; q < n; ++q) {
      const long i = d.cell[3 * q], j = d.cell[3 * q + 1], k = d.cell[3 * q + 2];
      if (i < 0 || i >= g.L || j < 0 || j >= g.ny || k < 0 || k >= g.nz) continue;
      key[q] = g.cell(i, j, k);
      order.push_back(q);
    }
    std::stable_sort(order.begin(), order.end(), [&](long a, long b) { return key[a] < key[b]; });
  }
  std::vector<int64_t> npart(nsorts, 0);
  for (long q : order) npart[1 + d.bunch[q]]++;

  // ---- particles file ----
  {
    h5::Writer w;
    for (int s = 1; s < nsorts; ++s) {
      if (!npart[s]) continue;
      char nm[64];
      snprintf(nm, sizeof nm, "Ions_of_sort_%2.2d_at_PE_0", s);
      if (!w.add(nm, h5::SAVED_ION, {(uint64_t)npart[s]})) return fail("save: " + w.error());
    }
    if (!w.add("Nparticles_at_PE_0", h5::I64, {(uint64_t)nsorts})) return fail("save: " + w.error());
    const std::string path = file_name(dir, isave, "particles");
    if (!w.open(path)) return fail("save: " + w.error());
    // sets are sorted by name: "Ions_of_sort_01..." < ... < "Nparticles_at_PE_0"
    int set = 0;
    for (int s = 1; s < nsorts; ++s) {
      if (!npart[s]) continue;
      std::vector<qss_saved_ion> rec;
      rec.reserve(npart[s]);
      for (long q : order) {
        if (1 + d.bunch[q] != s) continue;
        const long i = d.cell[3 * q], j = d.cell[3 * q + 1], k = d.cell[3 * q + 2];
        qss_saved_ion r{};
        r.cell = (int32_t)g.cell(i, j, k);
        r.x = (d.frac[3 * q] + (double)i) * c->p.hx;
        r.y = (d.frac[3 * q + 1] + (double)j) * c->hy;
        r.z = (d.frac[3 * q + 2] + (double)k) * c->hz;
        r.px = d.mom[3 * q]; r.py = d.mom[3 * q + 1]; r.pz = d.mom[3 * q + 2];
        r.weight = d.w[q];
        r.q2m = c->drv.par.q2m[s - 1];
        r.i_z = 1;
        r.sort = s;
        rec.push_back(r);
      }
      if (!w.put(set++, 0, rec.data(), rec.size() * sizeof(qss_saved_ion)))
        return fail("save: " + w.error());
    }
    if (!w.put(set, 0, npart.data(), npart.size() * sizeof(int64_t)) || !w.close())
      return fail("save: " + w.error());
  }

  // ---- fields file ----
  h5::Writer w;
  const uint64_t L = g.L, ny = g.ny, nz = g.nz;
  std::vector<std::string> names = {"meshinfo", "phase", "X", "Y", "Z"};
  if (!w.add("meshinfo", h5::I32, {16}) || !w.add("phase", h5::F64, {1}) ||
      !w.add("X", h5::F64, {L}) || !w.add("Y", h5::F64, {ny}) || !w.add("Z", h5::F64, {nz}))
    return fail("save: " + w.error());
  for (int f = 0; f < 10; ++f) {
    if (!w.add(kFieldName[f], h5::F64, {nz, ny, L})) return fail("save: " + w.error());
    names.push_back(kFieldName[f]);
  }
  for (int s = 0; s < nsorts; ++s) {
    const std::string nm = "Dens" + std::to_string(s);
    if (!w.add(nm, h5::F64, {nz, ny, L})) return fail("save: " + w.error());
    names.push_back(nm);
  }
  const std::string path = file_name(dir, isave, "fields");
  if (!w.open(path)) return fail("save: " + w.error());
  std::vector<std::string> sorted = names;
  std::sort(sorted.begin(), sorted.end(),
            [](const std::string &a, const std::string &b) { return strcmp(a.c_str(), b.c_str()) < 0; });
  auto idx = [&](const std::string &nm) {
    return (int)(std::find(sorted.begin(), sorted.end(), nm) - sorted.begin());
  };
  const long sxy = g.sx() * g.sy(), sxz = g.sx() * g.sz();
  const int32_t mi[16] = {(int32_t)g.L, kGhost, (int32_t)g.sx(), 0,
                          (int32_t)g.ny, kGhost, (int32_t)g.sy(), 0,
                          (int32_t)g.nz, kGhost, (int32_t)g.sz(), 0,
                          (int32_t)(sxy & 0xffffffff), (int32_t)(sxy >> 32),
                          (int32_t)(sxz & 0xffffffff), (int32_t)(sxz >> 32)};
  std::vector<double> ax(L), ay(ny), az(nz);
  for (uint64_t i = 0; i < L; ++i) ax[i] = ((double)i + shift) * c->p.hx;
  for (uint64_t j = 0; j < ny; ++j) ay[j] = (double)j * c->hy - c->p.ly / 2.;
  for (uint64_t k = 0; k < nz; ++k) az[k] = (double)k * c->hz - c->p.lz / 2.;
  if (!w.put(idx("meshinfo"), 0, mi, sizeof mi) || !w.put(idx("phase"), 0, &phase, 8) ||
      !w.put(idx("X"), 0, ax.data(), L * 8) || !w.put(idx("Y"), 0, ay.data(), ny * 8) ||
      !w.put(idx("Z"), 0, az.data(), nz * 8))
    return fail("save: " + w.error());
  std::vector<double> h;
  for (int f = 0; f < 9; ++f) {
    CHECK(fetch_history(c, kFieldId[f], h));
    if (!put_field(w, idx(kFieldName[f]), g, h)) return fail("save: " + w.error());
  }
  const size_t nc3 = (size_t)L * ny * nz;
  std::vector<double> zero;
  if (evolving) {
    // Mesh::Density(ALLSORTS): per particle of cell (i,j,k), fractions (x,y,z),
    // the 8 cells (i-a, j-b, k-c), a, b, c in {0,1}, get weight products of
    // x or 1-x etc. (Cell3Dm::AddDensity), charge weight*polarity into dens
    // and |weight*polarity| into Dens<sort>
    std::vector<double> dens(nc3, 0.0);
    std::vector<std::vector<double>> ds(nsorts);
    for (int s = 1; s < nsorts; ++s) ds[s].assign(nc3, 0.0);
    for (long q : order) {
      const long i = d.cell[3 * q], j = d.cell[3 * q + 1], k = d.cell[3 * q + 2];
      const double x = d.frac[3 * q], y = d.frac[3 * q + 1], z = d.frac[3 * q + 2];
      const double wq = d.w[q] * c->p.bunch[d.bunch[q]].polarity, aw = std::fabs(wq);
      std::vector<double> &dsort = ds[1 + d.bunch[q]];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int e = 0; e < 2; ++e) {
            const long ii = i - a, jj = j - b, kk = k - e;
            if (ii < 0 || jj < 0 || kk < 0) continue;
            const double f = (a ? 1. - x : x) * (b ? 1. - y : y) * (e ? 1. - z : z);
            const size_t m = (size_t)ii + L * ((size_t)jj + ny * (size_t)kk);
            dens[m] += f * wq;
            dsort[m] += f * aw;
          }
    }
    if (!put_cells(w, idx("dens"), dens)) return fail("save: " + w.error());
    zero.assign(nc3, 0.0);
    if (!put_cells(w, idx("Dens0"), zero)) return fail("save: " + w.error());
    for (int s = 1; s < nsorts; ++s)
      if (!put_cells(w, idx("Dens" + std::to_string(s)), ds[s])) return fail("save: " + w.error());
  } else {
    CHECK(fetch_history(c, QSS_RHOBEAM, h));
    if (!put_field(w, idx("dens"), g, h)) return fail("save: " + w.error());
    for (double &v : h) v = std::fabs(v);
    if (!put_field(w, idx("Dens1"), g, h)) return fail("save: " + w.error());
    zero.assign(nc3, 0.0);
    if (!put_cells(w, idx("Dens0"), zero)) return fail("save: " + w.error());
  }
  if (!w.close()) return fail("save: " + w.error());
  return QSS_OK;
}

extern "C" int qss_load(qss_ctx *c, const char *dir, int isave, double *phase) {
  if (!c) return fail("load: null context");
  if (isave < 0 || isave > 999) return fail("load: isave must be in [0, 999]");
  const Grid g{c->p.nslices, c->ny, c->nz};
  const std::string fpath = file_name(dir, isave, "fields");
  h5::Reader r;
  if (!r.open(fpath)) return fail("load: " + r.error());
  const h5::Dataset *mi = r.find("meshinfo");
  int32_t m[16] = {};
  if (!mi || mi->count() < 12 || !r.read(*mi, h5::I32, 0, 12, m))
    return fail("load: " + fpath + " has no readable meshinfo");
  if (m[0] != g.L || m[4] != g.ny || m[8] != g.nz)
    return fail("load: " + fpath + " holds a " + std::to_string(m[0]) + " x " +
                std::to_string(m[4]) + " x " + std::to_string(m[8]) + " mesh, the context " +
                std::to_string(g.L) + " x " + std::to_string(g.ny) + " x " + std::to_string(g.nz));
  if (phase) {
    const h5::Dataset *ph = r.find("phase");
    *phase = 0;
    if (ph && ph->count() >= 1 && !r.read(*ph, h5::F64, 0, 1, phase)) return fail("load: " + r.error());
  }
  // every field the context keeps history of
  const size_t nc = (size_t)c->ncell, n = (size_t)g.L * nc;
  std::vector<double> v((size_t)g.ny * g.L), h;
  for (int f = 0; f < 9; ++f) {
    const int id = kFieldId[f];
    if (!c->hist[id]) continue;
    const h5::Dataset *ds = r.find(kFieldName[f]);
    if (!ds || ds->dims.size() != 3 || ds->dims[0] != (uint64_t)g.nz ||
        ds->dims[1] != (uint64_t)g.ny || ds->dims[2] != (uint64_t)g.L)
      return fail(std::string("load: ") + fpath + " has no " + kFieldName[f] +
                  " dataset of the context's shape");
    h.assign(n, 0.0);
    for (long k = 0; k < g.nz; ++k) {
      if (!r.read(*ds, h5::F64, (uint64_t)k * v.size(), v.size(), v.data()))
        return fail("load: " + r.error());
      for (long j = 0; j < g.ny; ++j)
        for (long i = 0; i < g.L; ++i)
          h[(size_t)(g.L - 1 - i) * nc + j + (size_t)g.ny * k] = v[(size_t)j * g.L + i];
    }
    HIPCHECK(hipSetDevice(c->device));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(c->hist[id] + (size_t)c->sel * n, h.data(), n * sizeof(double),
                       hipMemcpyHostToDevice));
  }
  if (!c->drv.on) return QSS_OK;   // particles go into a seeded evolving driver

  const std::string ppath = file_name(dir, isave, "particles");
  h5::Reader pr;
  if (!pr.open(ppath)) return fail("load: " + pr.error());
  std::vector<int32_t> cell, bunch;
  std::vector<double> frac, mom, w;
  for (int s = 1; s <= c->p.nbunch; ++s) {
    char nm[64];
    snprintf(nm, sizeof nm, "Ions_of_sort_%2.2d_at_PE_0", s);
    const h5::Dataset *ds = pr.find(nm);
    if (!ds) continue;
    if (ds->dims.size() != 1) return fail(std::string("load: ") + nm + " is not a 1-D record set");
    std::vector<qss_saved_ion> rec(ds->count());
    if (!pr.read(*ds, h5::SAVED_ION, 0, rec.size(), rec.data())) return fail("load: " + pr.error());
    for (const qss_saved_ion &p : rec) {
      const long cn = p.cell;
      long i = cn % g.sx(), rest = cn / g.sx();
      long j = rest % g.sy(), k = rest / g.sy();
      i -= kGhost; j -= kGhost; k -= kGhost;
      if (cn < 0 || i < -1 || i > g.L || j < -1 || j > g.ny || k < -1 || k > g.nz)
        return fail(std::string("load: ") + nm + ": cell number " + std::to_string(cn) +
                    " outside the mesh");
      if (p.sort != s) return fail(std::string("load: ") + nm + ": record of another sort");
      const double f3[3] = {p.x / c->p.hx - (double)i, p.y / c->hy - (double)j,
                            p.z / c->hz - (double)k};
      cell.push_back((int32_t)i); cell.push_back((int32_t)j); cell.push_back((int32_t)k);
      for (double f : f3) frac.push_back(std::min(1.0, std::max(0.0, f)));
      mom.push_back(p.px); mom.push_back(p.py); mom.push_back(p.pz);
      w.push_back(p.weight);
      bunch.push_back(s - 1);
    }
  }
  return qss_driver_set(c, (int64_t)w.size(), cell.data(), frac.data(), mom.data(), w.data(),
                        bunch.data());
}
