// Plan lifecycle: grid constants, paired transform matrices, modal tables.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "plan.h"

namespace nfgs {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

// Exact-argument type-I matrices: the integer phase is reduced before the
// trigonometric call so every entry is correctly rounded in double.  Entries
// depend on the phase only, so the trigonometric values of one period are
// tabulated once per matrix.
class Type1 {
 public:
  Type1(int kind, int n) : dct_(kind == NFGS_DCT1), n_(n) {
    const double pi = 3.14159265358979323846;
    period_ = dct_ ? 2L * (n - 1) : 2L * (n + 1);
    tab_.resize((size_t)period_);
    for (long ph = 0; ph < period_; ++ph)
      tab_[(size_t)ph] = dct_ ? std::cos(pi * (double)ph / (double)(n - 1))
                              : 2.0 * std::sin(pi * (double)ph / (double)(n + 1));
  }
  double operator()(int k, int j) const {
    if (dct_) {
      const double w = (j == 0 || j == n_ - 1) ? 1.0 : 2.0;
      return w * tab_[(size_t)(((long)j * k) % period_)];
    }
    return tab_[(size_t)(((long)(j + 1) * (k + 1)) % period_)];
  }

 private:
  bool dct_;
  int n_;
  long period_;
  std::vector<double> tab_;
};

template <typename T>
int upload(const std::vector<double>& host, void** dev) {
  std::vector<T> cast(host.size());
  for (size_t i = 0; i < host.size(); ++i) cast[i] = static_cast<T>(host[i]);
  NFGS_HIP_CHECK(hipMalloc(dev, cast.size() * sizeof(T)));
  NFGS_HIP_CHECK(hipMemcpy(*dev, cast.data(), cast.size() * sizeof(T), hipMemcpyHostToDevice));
  return NFGS_OK;
}

int upload_dtype(int dtype, const std::vector<double>& host, void** dev) {
  return dtype == NFGS_F32 ? upload<float>(host, dev) : upload<double>(host, dev);
}

// Digit form of E and O for the exact-integer FP32 transform.  Entry (k, j')
// of a half becomes the integer round(T[k][2j'+half] * 2^shift) (|.| <= 2^30)
// written as four signed base-256 digits; digit p of lane l of 16-row block b,
// K chunk c, holds in byte 4w+i the paired row j' = 4(l/16) + w + 16(4c+i)
// of output row k = 16b + l%16 (the order in which the kernel packs the
// vector digits).  Paired rows past 64*nch (at most one) form the FP64 tail;
// when h = 16*hb + 1 the last output row is kept whole (mrem) in FP64.
int build_digits(int kind, int n, const Type1& T1, PairedMatrix* m) {
  const int h = m->h, ne = m->rows;
  const int nch = std::max(1, (ne + 62) / 64);
  const int rem = (h > 16 && h % 16 == 1) ? 1 : 0;
  const int hb = rem ? h / 16 : (h + 15) / 16;
  if (nch > 2 || hb > 8) return NFGS_OK;  // longer vectors keep the FP64 kernel
  m->nch = nch;
  m->tail = ne > 64 * nch ? 1 : 0;
  m->hb = hb;
  m->rem = rem;
  const double inv_scale = kind == NFGS_DCT1 ? 1.0 / (2.0 * (n - 1)) : 1.0 / (2.0 * (n + 1));
  for (int dir = 0; dir < 2; ++dir) {
    const double sc = dir == NFGS_FORWARD ? 1.0 : inv_scale;
    auto entry = [&](int half, int k, int jp) -> double {
      const int j = 2 * jp + half;
      return (k < h && j < n) ? sc * T1(k, j) : 0.0;
    };
    double mx = 0.0;
    for (int half = 0; half < 2; ++half)
      for (int k = 0; k < h; ++k)
        for (int jp = 0; jp < ne; ++jp) mx = std::max(mx, std::fabs(entry(half, k, jp)));
    int e = 0;
    (void)std::frexp(mx > 0.0 ? mx : 1.0, &e);  // mx < 2^e
    int sh = 30 - e;
    if (std::ldexp(mx, sh + 1) <= 1073741824.0) ++sh;
    m->shift[dir] = sh;
    // biased digit words of every (half, k, j') the blocks hold
    const int kr = 16 * hb, jr = 64 * nch;
    std::vector<uint32_t> words((size_t)2 * kr * jr);
    for (int half = 0; half < 2; ++half)
      for (int k = 0; k < kr; ++k)
        for (int jp = 0; jp < jr; ++jp) {
          const long mi = std::lrint(std::ldexp(entry(half, k, jp), sh));
          words[((size_t)half * kr + k) * jr + jp] = ((uint32_t)(int32_t)mi + 0x808080u) ^ 0x808080u;
        }
    const size_t nd = (size_t)2 * 4 * nch * hb * 64 * 16;
    std::vector<int8_t> dg(nd);
    size_t idx = 0;
    for (int half = 0; half < 2; ++half)
      for (int p = 0; p < 4; ++p)
        for (int c = 0; c < nch; ++c)
          for (int b = 0; b < hb; ++b)
            for (int l = 0; l < 64; ++l)
              for (int byte = 0; byte < 16; ++byte) {
                const int w = byte >> 2, i = byte & 3;
                const int k = 16 * b + (l & 15);
                const int jp = 4 * (l >> 4) + w + 16 * (4 * c + i);
                const uint32_t d = words[((size_t)half * kr + k) * jr + jp];
                dg[idx++] = (int8_t)(uint8_t)(d >> (8 * (3 - p)));
              }
    NFGS_HIP_CHECK(hipMalloc(&m->dig[dir], nd));
    NFGS_HIP_CHECK(hipMemcpy(m->dig[dir], dg.data(), nd, hipMemcpyHostToDevice));
    std::vector<double> mt((size_t)2 * 16 * hb, 0.0);
    if (m->tail)
      for (int half = 0; half < 2; ++half)
        for (int k = 0; k < 16 * hb; ++k) mt[(size_t)half * 16 * hb + k] = entry(half, k, 64 * nch);
    int rc = upload<double>(mt, &m->mtail[dir]);
    if (rc) return rc;
    std::vector<double> mr((size_t)n, 0.0);
    if (rem)
      for (int j = 0; j < n; ++j) mr[j] = sc * T1(16 * hb, j);
    rc = upload<double>(mr, &m->mrem[dir]);
    if (rc) return rc;
  }
  return NFGS_OK;
}

// bf16 of a value, round to nearest even (finite values)
uint16_t bf16_of(double v) {
  const float f = static_cast<float>(v);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
double bf16_value(uint16_t h) {
  const uint32_t u = static_cast<uint32_t>(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Split-bf16 operand fragments of E and O (PairedMatrix::bfr) for one
// direction: entry (K, out) of E or O as hi + lo, hi = bf16(t), lo = bf16(t - hi).
int build_bf16(const std::vector<double>& et, const std::vector<double>& ot, int dir,
               PairedMatrix* m) {
  const int nc = (m->rows + 15) / 16, nb = (m->h + 15) / 16;
  m->bchunks = nc;
  m->bblocks = nb;
  std::vector<uint16_t> f((size_t)4 * nc * nb * 64 * 4, 0);
  for (int eo = 0; eo < 2; ++eo) {
    const std::vector<double>& t = eo == 0 ? et : ot;
    for (int c = 0; c < nc; ++c)
      for (int b = 0; b < nb; ++b)
        for (int l = 0; l < 64; ++l)
          for (int q = 0; q < 4; ++q) {
            const int K = c * 16 + 4 * (l / 16) + q, o = b * 16 + (l % 16);
            const double v = (K < m->rows && o < m->h) ? t[(size_t)K * m->cols_pad + o] : 0.0;
            const uint16_t hi = bf16_of(v);
            const uint16_t lo = bf16_of(v - bf16_value(hi));
            // fragment index ((part * nc + c) * nb + b) * 64 + l, 4 bf16 each
            const size_t base = ((size_t)c * nb + b) * 64 + l, part = (size_t)nc * nb * 64;
            f[((size_t)(2 * eo) * part + base) * 4 + q] = hi;
            f[((size_t)(2 * eo + 1) * part + base) * 4 + q] = lo;
          }
  }
  NFGS_HIP_CHECK(hipMalloc(&m->bfr[dir], f.size() * sizeof(uint16_t)));
  NFGS_HIP_CHECK(hipMemcpy(m->bfr[dir], f.data(), f.size() * sizeof(uint16_t),
                           hipMemcpyHostToDevice));
  return NFGS_OK;
}

int build_paired(int dtype, int kind, int n, PairedMatrix* m) {
  m->n = n;
  m->h = (n + 1) / 2;
  m->rows = (n + 1) / 2;
  m->rows_pad = ((m->rows + 15) / 16) * 16;
  m->cols_pad = ((m->h + 15) / 16) * 16;
  const double inv_scale = kind == NFGS_DCT1 ? 1.0 / (2.0 * (n - 1)) : 1.0 / (2.0 * (n + 1));
  const Type1 T1(kind, n);
  for (int dir = 0; dir < 2; ++dir) {
    const double sc = dir == NFGS_FORWARD ? 1.0 : inv_scale;
    std::vector<double> et((size_t)m->rows_pad * m->cols_pad, 0.0);
    std::vector<double> ot((size_t)m->rows_pad * m->cols_pad, 0.0);
    for (int jp = 0; jp < m->rows; ++jp) {
      for (int k = 0; k < m->h; ++k) {
        const int je = 2 * jp, jo = 2 * jp + 1;
        if (je < n) et[(size_t)jp * m->cols_pad + k] = sc * T1(k, je);
        if (jo < n) ot[(size_t)jp * m->cols_pad + k] = sc * T1(k, jo);
      }
    }
    // FP64 for every plan dtype: FP64 plans (and FP32 vectors too long for
    // the digit form) run their products in FP64
    int rc = upload<double>(et, &m->et[dir]);
    if (rc) return rc;
    rc = upload<double>(ot, &m->ot[dir]);
    if (rc) return rc;
    if (dtype == NFGS_F32) {
      rc = build_bf16(et, ot, dir, m);
      if (rc) return rc;
    }
  }
  return dtype == NFGS_F32 ? build_digits(kind, n, T1, m) : NFGS_OK;
}

void free_paired(PairedMatrix* m) {
  for (int d = 0; d < 2; ++d) {
    if (m->et[d]) (void)hipFree(m->et[d]);
    if (m->ot[d]) (void)hipFree(m->ot[d]);
    if (m->dig[d]) (void)hipFree(m->dig[d]);
    if (m->mtail[d]) (void)hipFree(m->mtail[d]);
    if (m->mrem[d]) (void)hipFree(m->mrem[d]);
    if (m->bfr[d]) (void)hipFree(m->bfr[d]);
    m->et[d] = m->ot[d] = m->dig[d] = m->mtail[d] = m->mrem[d] = m->bfr[d] = nullptr;
  }
}

}  // namespace

const PairedMatrix* find_matrix(const nfgs_plan* p, int kind, int n) {
  if (kind == NFGS_DCT1) {
    if (n == p->dct_x.n) return &p->dct_x;
    if (n == p->dct_y.n) return &p->dct_y;
  } else if (kind == NFGS_DST1) {
    if (n == p->dst_x.n) return &p->dst_x;
    if (n == p->dst_y.n) return &p->dst_y;
  }
  return nullptr;
}

}  // namespace nfgs

extern "C" const char* nfgs_last_error(void) { return nfgs::g_last_error.c_str(); }

// SHA-256 of the sources the library was compiled from (set by the build).
#ifndef NFGS_SOURCE_DIGEST
#define NFGS_SOURCE_DIGEST "unknown"
#endif
extern "C" const char* nfgs_source_digest(void) { return NFGS_SOURCE_DIGEST; }

extern "C" int nfgs_plan_destroy(nfgs_plan* p) {
  if (!p) return NFGS_OK;
  nfgs::free_paired(&p->dct_x);
  nfgs::free_paired(&p->dct_y);
  nfgs::free_paired(&p->dst_x);
  nfgs::free_paired(&p->dst_y);
  if (p->lam) (void)hipFree(p->lam);
  if (p->kk) (void)hipFree(p->kk);
  if (p->lamd) (void)hipFree(p->lamd);
  if (p->ax) (void)hipFree(p->ax);
  if (p->ay) (void)hipFree(p->ay);
  if (p->red) (void)hipFree(p->red);
  delete p;
  return NFGS_OK;
}

extern "C" int nfgs_plan_create(const nfgs_grid_desc* d, nfgs_plan** out) {
  using namespace nfgs;
  NFGS_REQUIRE(d != nullptr && out != nullptr, NFGS_ERR_ARG, "plan_create: null argument");
  *out = nullptr;
  NFGS_REQUIRE(d->nx >= 3 && d->ny >= 3 && d->nz >= 3, NFGS_ERR_ARG,
               "plan_create: every calculation-grid dimension must be >= 3");
  NFGS_REQUIRE(d->dtype == NFGS_F32 || d->dtype == NFGS_F64, NFGS_ERR_ARG,
               "plan_create: dtype must be NFGS_F32 or NFGS_F64");
  NFGS_REQUIRE(d->dx > 0 && d->dy > 0 && d->dz > 0, NFGS_ERR_ARG,
               "plan_create: grid spacings must be positive");
  NFGS_REQUIRE(d->rho >= 0.0 && d->rho <= 1.0, NFGS_ERR_ARG,
               "bottom_divergence_robin_weight must be between 0 and 1");
  NFGS_REQUIRE(d->win_z0 >= 0 && d->win_z1 <= d->nz - 1 && d->win_y0 >= 0 &&
                   d->win_y1 <= d->ny && d->win_x0 >= 0 && d->win_x1 <= d->nx,
               NFGS_ERR_ARG, "plan_create: sigma window outside the face grid");
  NFGS_REQUIRE((long)d->nx * d->ny * d->nz < (1L << 40), NFGS_ERR_ARG, "plan_create: grid too large");

  nfgs_plan* p = new nfgs_plan();
  p->d = *d;
  p->plane = (long)d->ny * d->nx;
  p->iplane = (long)(d->ny - 2) * (d->nx - 2);
  p->N = p->plane * d->nz;
  p->esize = d->dtype == NFGS_F32 ? 4 : 8;
  // Partials are laid out for the largest block count any admissible
  // z-chunking launches (the cost model's or a tuned count's).
  p->pblocks = (int)shape::partial_capacity(d->nx, d->ny, d->nz, cu_count_stencil());

  int rc = build_paired(d->dtype, NFGS_DCT1, d->nx, &p->dct_x);
  if (!rc) rc = build_paired(d->dtype, NFGS_DCT1, d->ny, &p->dct_y);
  if (!rc) rc = build_paired(d->dtype, NFGS_DST1, d->nx - 2, &p->dst_x);
  if (!rc) rc = build_paired(d->dtype, NFGS_DST1, d->ny - 2, &p->dst_y);

  if (!rc) {
    // Endpoint-Neumann eigenvalues (nfgs_solver.py:1098-1109) and k = |lambda|^(1/2).
    const double pi = 3.14159265358979323846;
    std::vector<double> lam((size_t)p->plane), kk((size_t)p->plane);
    std::vector<double> lx(d->nx), ly(d->ny);
    for (int i = 0; i < d->nx; ++i) {
      const double s = std::sin(pi * i / (2.0 * (d->nx - 1)));
      lx[i] = -4.0 * s * s / (d->dx * d->dx);
    }
    for (int j = 0; j < d->ny; ++j) {
      const double s = std::sin(pi * j / (2.0 * (d->ny - 1)));
      ly[j] = -4.0 * s * s / (d->dy * d->dy);
    }
    for (int j = 0; j < d->ny; ++j)
      for (int i = 0; i < d->nx; ++i) {
        const double v = lx[i] + ly[j];
        lam[(size_t)j * d->nx + i] = v;
        kk[(size_t)j * d->nx + i] = std::sqrt(v < 0 ? -v : 0.0);
      }
    // Endpoint-Dirichlet interior eigenvalues (nfgs_solver.py:1083-1096).
    std::vector<double> lamd((size_t)p->iplane);
    const int nxi = d->nx - 2, nyi = d->ny - 2;
    std::vector<double> dx_(nxi), dy_(nyi);
    for (int i = 0; i < nxi; ++i) {
      const double s = std::sin(pi * (i + 1) / (2.0 * (nxi + 1)));
      dx_[i] = -4.0 * s * s / (d->dx * d->dx);
    }
    for (int j = 0; j < nyi; ++j) {
      const double s = std::sin(pi * (j + 1) / (2.0 * (nyi + 1)));
      dy_[j] = -4.0 * s * s / (d->dy * d->dy);
    }
    for (int j = 0; j < nyi; ++j)
      for (int i = 0; i < nxi; ++i) lamd[(size_t)j * nxi + i] = dx_[i] + dy_[j];
    // |d/dx| factors of the spectral horizontal operator (nfgs_solver.py:4357-4383).
    std::vector<double> ax(d->nx), ay(d->ny);
    for (int i = 0; i < d->nx; ++i) ax[i] = 2.0 * std::sin(pi * i / (2.0 * (d->nx - 1))) / d->dx;
    for (int j = 0; j < d->ny; ++j) ay[j] = 2.0 * std::sin(pi * j / (2.0 * (d->ny - 1))) / d->dy;
    rc = upload_dtype(d->dtype, lam, &p->lam);
    if (!rc) rc = upload_dtype(d->dtype, ax, &p->ax);
    if (!rc) rc = upload_dtype(d->dtype, ay, &p->ay);
    if (!rc) rc = upload_dtype(d->dtype, kk, &p->kk);
    if (!rc) rc = upload_dtype(d->dtype, lamd, &p->lamd);
    if (!rc && hipMalloc(reinterpret_cast<void**>(&p->red), kRedScratch * sizeof(double)) !=
                   hipSuccess) {
      set_error("plan_create: hipMalloc of the reduction scratch failed");
      rc = NFGS_ERR_HIP;
    }
  }
  if (rc) {
    nfgs_plan_destroy(p);
    return rc == NFGS_ERR_HIP ? NFGS_ERR_ALLOC : rc;
  }
  *out = p;
  return NFGS_OK;
}

extern "C" long nfgs_partials_count(const nfgs_plan* p) {
  return p ? (long)p->pblocks * NFGS_NQ : 0;
}

extern "C" const void* nfgs_plan_table(const nfgs_plan* p, int which) {
  if (!p) return nullptr;
  switch (which) {
    case 0: return p->lam;
    case 1: return p->kk;
    case 2: return p->lamd;
    case 3: return p->ax;
    case 4: return p->ay;
    default: return nullptr;
  }
}

extern "C" long nfgs_plan_table_size(const nfgs_plan* p, int which) {
  if (!p) return 0;
  switch (which) {
    case 0:
    case 1: return p->plane;
    case 2: return p->iplane;
    case 3: return p->d.nx;
    case 4: return p->d.ny;
    default: return 0;
  }
}

extern "C" int nfgs_copy_plan_table(const nfgs_plan* p, int which, void* dst, void* stream) {
  NFGS_REQUIRE(p && dst, NFGS_ERR_ARG, "copy_plan_table: null argument");
  const void* src = nfgs_plan_table(p, which);
  NFGS_REQUIRE(src, NFGS_ERR_ARG, "copy_plan_table: no such table");
  NFGS_HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)nfgs_plan_table_size(p, which) * p->esize,
                                hipMemcpyDeviceToDevice, nfgs::as_stream(stream)));
  return NFGS_OK;
}

extern "C" int nfgs_plan_set_tuning(nfgs_plan* p, int knob, int value) {
  NFGS_REQUIRE(p, NFGS_ERR_ARG, "plan_set_tuning: null plan");
  NFGS_REQUIRE(value >= 0, NFGS_ERR_ARG, "plan_set_tuning: value must be >= 0");
  if (knob == NFGS_TUNE_ZCHUNKS) {
    NFGS_REQUIRE(value == 0 || nfgs::shape::zchunks_fit(p->d.nz, value), NFGS_ERR_ARG,
                 "plan_set_tuning: z-chunk count does not fit the grid");
    const int old = p->zchunks;
    p->zchunks = value;
    if (nfgs::tiled_blocks(p) > p->pblocks) {  // not for counts zchunks_fit admits
      p->zchunks = old;
      nfgs::set_error("plan_set_tuning: more blocks than the partials hold");
      return NFGS_ERR_ARG;
    }
    return NFGS_OK;
  }
  if (knob == NFGS_TUNE_TRANSFORM_BLOCKS_PER_CU) {
    for (PairedMatrix* m : {&p->dct_x, &p->dct_y, &p->dst_x, &p->dst_y}) m->blocks_per_cu = value;
    return NFGS_OK;
  }
  nfgs::set_error("plan_set_tuning: unknown knob");
  return NFGS_ERR_ARG;
}

extern "C" int nfgs_plan_set_lateral_arithmetic(nfgs_plan* p, int arith) {
  NFGS_REQUIRE(p, NFGS_ERR_ARG, "plan_set_lateral_arithmetic: null plan");
  NFGS_REQUIRE(arith == NFGS_LATERAL_EXACT || arith == NFGS_LATERAL_BF16X3, NFGS_ERR_ARG,
               "plan_set_lateral_arithmetic: arithmetic must be NFGS_LATERAL_EXACT or "
               "NFGS_LATERAL_BF16X3");
  NFGS_REQUIRE(arith == NFGS_LATERAL_EXACT || p->d.dtype == NFGS_F32, NFGS_ERR_ARG,
               "plan_set_lateral_arithmetic: the split-bf16 products are for FP32 plans");
  p->lateral_arith = arith;
  return NFGS_OK;
}

extern "C" int nfgs_plan_get_lateral_arithmetic(const nfgs_plan* p) {
  return p ? p->lateral_arith : -1;
}

extern "C" int nfgs_plan_get_tuning(const nfgs_plan* p, int knob) {
  if (!p) return -1;
  if (knob == NFGS_TUNE_ZCHUNKS) return p->zchunks;
  if (knob == NFGS_TUNE_TRANSFORM_BLOCKS_PER_CU) return p->dct_x.blocks_per_cu;
  return -1;
}
