// CPU ORACLE — C entry points for the parity tests (test infrastructure only).
// Built by oracle/Makefile into oracle/_build/liboracle.so; loaded only by tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg.
#include "oracle.hpp"
#include "oracle_cone.hpp"
#include "oracle_dos.hpp"
#include "oracle_lc.hpp"
#include "oracle_grad.hpp"
#include "oracle_scale.hpp"
#include <omp.h>

using namespace orc;

namespace {
Volume mkvol(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz) {
  Volume v;
  v.nx = nx; v.ny = ny; v.nz = nz; v.dtype = dtype; v.data = data; v.sx = sx; v.sy = sy; v.sz = sz;
  return v;
}
CamIn mkcam(const float* c) {
  CamIn k;
  for (int i = 0; i < 3; ++i) { k.eye[i] = c[i]; k.center[i] = c[3 + i]; k.up[i] = c[6 + i]; }
  k.fovy_deg = c[9];
  k.aspect = c[10];
  return k;
}
void threads(int n) { if (n > 0) omp_set_num_threads(n); }
}  // namespace

extern "C" {

int orc_max_threads() { return omp_get_max_threads(); }

void* orc_tf_new(int max_density, int ext_input, int n_rgb, const double* rgb_iso, int n_a, const double* a_iso) {
  TF* tf = new TF();
  if (!tf_build(*tf, max_density, ext_input != 0, n_rgb, rgb_iso, n_a, a_iso)) {
    delete tf;
    return nullptr;
  }
  return tf;
}
void orc_tf_free(void* tf) { delete static_cast<TF*>(tf); }
int orc_tf_size(void* tf) { return static_cast<TF*>(tf)->size(); }
void orc_tf_tables(void* h, double* tab, float* rgbt, float* rgba) {
  TF* tf = static_cast<TF*>(h);
  if (tab) std::memcpy(tab, tf->tab.data(), tf->tab.size() * sizeof(double));
  if (rgbt) std::memcpy(rgbt, tf->rgbt.data(), tf->rgbt.size() * sizeof(float));
  if (rgba) std::memcpy(rgba, tf->rgba.data(), tf->rgba.size() * sizeof(float));
}
void orc_tf_get(void* h, double value, double maxv, float* out) { tf_get(*static_cast<TF*>(h), value, maxv, out); }
float orc_tf_get_ext_n(void* h, double v) { return tf_get_ext_n(*static_cast<TF*>(h), v); }

void orc_tex1d(const float* table, int T, double d, double* out) { tex1d_tf<double>(table, T, d, out); }
double orc_tex3d_volume(int nx, int ny, int nz, int dtype, const void* data, double u, double v, double w) {
  Volume vol = mkvol(nx, ny, nz, dtype, data, 1, 1, 1);
  return tex3d_volume<double>(vol, V3<double>(u, v, w));
}
double orc_tex3d_grid(const double* g, int nx, int ny, int nz, double u, double v, double w) {
  return tex3d_grid<double, double>(g, nx, ny, nz, V3<double>(u, v, w));
}

// Per-pixel ray: dir xyz, tnear, tfar, hit (8 doubles, the last two unused).
void orc_rays(const float* cam, int W, int H, double gx, double gy, double gz, double* out) {
  CamIn c = mkcam(cam);
  CamBasis<double> cb = cam_basis<double>(c, W, H);
  V3<double> G(gx, gy, gz);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      V3<double> d = ray_dir(cb, x, y, W, H);
      double tn, tf;
      bool hit = ray_aabb(cb.eye, d, G, tn, tf);
      double* o = out + (size_t(y) * W + x) * 8;
      o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = tn; o[4] = tf; o[5] = hit ? 1 : 0; o[6] = o[7] = 0;
    }
}

void orc_render_rc1p_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                         void* tf, const float* cam, int W, int H, double step, double et, double* out,
                         double* margin, int nthreads, int stride) {
  threads(nthreads);
  render_rc1p<double>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), mkcam(cam), W, H, step, et,
                      out, margin, stride);
}
void orc_render_rc1p_f32(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                         void* tf, const float* cam, int W, int H, float step, float et, float* out, int nthreads) {
  threads(nthreads);
  render_rc1p<float>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), mkcam(cam), W, H, step, et,
                     out, nullptr);
}

// SAT over (nx+2)(ny+2)(nz+2) doubles, the reference's input grid + recurrence.
void orc_sat_f64(int nx, int ny, int nz, int dtype, const void* data, void* tf, double* out, int nthreads) {
  threads(nthreads);
  Volume v = mkvol(nx, ny, nz, dtype, data, 1, 1, 1);
  std::vector<double> g;
  sat_input(v, *static_cast<TF*>(tf), g);
  sat_build_recurrence(g, nx + 2, ny + 2, nz + 2);
  std::memcpy(out, g.data(), g.size() * sizeof(double));
}
// The per-voxel SAT input (border included) without the prefix sum.
void orc_sat_input_f64(int nx, int ny, int nz, int dtype, const void* data, void* tf, double* out) {
  Volume v = mkvol(nx, ny, nz, dtype, data, 1, 1, 1);
  std::vector<double> g;
  sat_input(v, *static_cast<TF*>(tf), g);
  std::memcpy(out, g.data(), g.size() * sizeof(double));
}

// EbsParams mirror for ctypes: ints then floats, see oracle.hpp.
struct EbsC {
  int apply_occlusion, apply_shadow, ao_shells, shadow_type;
  float ao_radius, sdw_angle_deg, sdw_interval, sdw_initial, sdw_weight, sdw_max_dist;
  float light_pos[3], light_fwd[3];
  float ka, kd;
};
static EbsParams ebs_from(const EbsC* c) {
  EbsParams p;
  p.apply_occlusion = c->apply_occlusion; p.apply_shadow = c->apply_shadow;
  p.ao_shells = c->ao_shells; p.shadow_type = c->shadow_type;
  p.ao_radius = c->ao_radius;
  p.sdw_angle_rad = float(double(c->sdw_angle_deg) * 3.14159265358979323846 / 180.0);
  p.sdw_interval = c->sdw_interval; p.sdw_initial = c->sdw_initial; p.sdw_weight = c->sdw_weight;
  p.sdw_max_dist = c->sdw_max_dist;
  for (int i = 0; i < 3; ++i) { p.light_pos[i] = c->light_pos[i]; p.light_fwd[i] = c->light_fwd[i]; }
  p.ka = c->ka; p.kd = c->kd;
  return p;
}
void orc_render_ebs_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                        void* tf, const double* sat, const EbsC* prm, const float* cam, int W, int H, double step,
                        double et, double* out, double* margin, int nthreads, int stride) {
  threads(nthreads);
  render_ebs<double, double>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), sat, ebs_from(prm),
                             mkcam(cam), W, H, step, et, out, margin, stride);
}
void orc_render_ebs_f32(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                        void* tf, const float* sat, const EbsC* prm, const float* cam, int W, int H, float step,
                        float et, float* out, int nthreads, int stride) {
  threads(nthreads);
  render_ebs<float, float>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), sat, ebs_from(prm),
                           mkcam(cam), W, H, step, et, out, nullptr, stride);
}

// AO and shadow factors of the EBS lighting at one texture-space position (FP64).
void orc_ebs_point(int nx, int ny, int nz, double sx, double sy, double sz, const double* sat, const EbsC* prm,
                   double px, double py, double pz, double* ao, double* sdw) {
  Volume v = mkvol(nx, ny, nz, 0, nullptr, sx, sy, sz);
  Ebs<double, double> e(sat, v, ebs_from(prm));
  V3<double> p(px, py, pz);
  *ao = e.ambient_occlusion(p);
  *sdw = e.shadows(p, nullptr);
}

// ------------------------------------------------------------------ pyramid / cones / MLP
struct ConeC {
  int n_steps, n_occ, apply_occ, apply_sdw;
  double occ_half_deg, ring_deg, sdw_half_deg, t0, occ_max, sdw_max, r_sigma, sigma0, ka, kd;
  double light_fwd[3];
};
static ConeParams cone_from(const ConeC* c) {
  ConeParams p;
  p.n_steps = c->n_steps; p.n_occ = c->n_occ; p.apply_occ = c->apply_occ; p.apply_sdw = c->apply_sdw;
  p.occ_half_deg = c->occ_half_deg; p.ring_deg = c->ring_deg; p.sdw_half_deg = c->sdw_half_deg;
  p.t0 = c->t0; p.occ_max = c->occ_max; p.sdw_max = c->sdw_max; p.r_sigma = c->r_sigma; p.sigma0 = c->sigma0;
  p.ka = c->ka; p.kd = c->kd;
  for (int i = 0; i < 3; ++i) p.light_fwd[i] = c->light_fwd[i];
  return p;
}

void* orc_pyr_new(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz, void* tf,
                  int bx, int by, int bz, double S0, int nlev, int nthreads) {
  threads(nthreads);
  Pyramid* P = new Pyramid();
  pyramid_build(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), bx, by, bz, S0, nlev, *P);
  return P;
}
// A pyramid made of given extinction levels (concatenated, level 0 first).
void* orc_pyr_wrap(double gx, double gy, double gz, int nlev, const int* dims, const double* tau) {
  Pyramid* P = new Pyramid();
  P->G[0] = gx; P->G[1] = gy; P->G[2] = gz;
  P->levels = nlev;
  size_t o = 0;
  for (int i = 0; i < nlev; ++i) {
    P->nx.push_back(dims[3 * i]); P->ny.push_back(dims[3 * i + 1]); P->nz.push_back(dims[3 * i + 2]);
    size_t n = size_t(dims[3 * i]) * dims[3 * i + 1] * dims[3 * i + 2];
    P->tau.emplace_back(tau + o, tau + o + n);
    o += n;
  }
  return P;
}
void orc_pyr_free(void* h) { delete static_cast<Pyramid*>(h); }
int orc_pyr_levels(void* h) { return static_cast<Pyramid*>(h)->levels; }
void orc_pyr_dims(void* h, int l, int* d) {
  Pyramid* P = static_cast<Pyramid*>(h);
  d[0] = P->nx[l]; d[1] = P->ny[l]; d[2] = P->nz[l];
}
void orc_pyr_copy(void* h, int l, double* tau, double* op) {
  Pyramid* P = static_cast<Pyramid*>(h);
  if (tau) std::memcpy(tau, P->tau[l].data(), P->tau[l].size() * sizeof(double));
  if (op && !P->op.empty()) std::memcpy(op, P->op[l].data(), P->op[l].size() * sizeof(double));
}
// One voxel of level 0 (opacity), for sparse checks of large pyramids.
double orc_pyr_level0_at(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                         void* tf, int bx, int by, int bz, double S0, int x, int y, int z) {
  Volume v = mkvol(nx, ny, nz, dtype, data, sx, sy, sz);
  const TF& t = *static_cast<TF*>(tf);
  double G[3] = {nx * sx, ny * sy, nz * sz};
  double vs[3] = {G[0] / bx, G[1] / by, G[2] / bz};
  double gp[3] = {(x + 0.5) * vs[0], (y + 0.5) * vs[1], (z + 0.5) * vs[2]};
  double swc = 0, sw = 0;
  for (int i = -3; i <= 3; ++i)
    for (int j = -3; j <= 3; ++j)
      for (int k = -3; k <= 3; ++k) {
        double f[3] = {i * S0, j * S0, k * S0};
        double wk = (S0 * S0 * S0) * std::exp(-(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) / (2.0 * S0 * S0));
        double tt[3] = {(gp[0] + f[0]) / G[0], (gp[1] + f[1]) / G[1], (gp[2] + f[2]) / G[2]};
        double ck = 0;
        if (!(tt[0] < 0 || tt[1] < 0 || tt[2] < 0 || tt[0] > 1 || tt[1] > 1 || tt[2] > 1)) {
          double d = tex3d_volume<double>(v, V3<double>(tt[0], tt[1], tt[2]));
          double c[4];
          tex1d_tf<double>(t.rgba.data(), t.size(), d, c);
          ck = c[3];
        }
        swc += wk * ck;
        sw += wk;
      }
  return swc / sw;
}
// One voxel of level i (opacity) from the given opacity level i-1.
double orc_pyr_next_at(const double* prev, int px, int py, int pz, int nx, int ny, int nz, double Si, double gx,
                       double gy, double gz, int x, int y, int z) {
  double G[3] = {gx, gy, gz};
  double lvs[3] = {G[0] / nx, G[1] / ny, G[2] / nz};
  double gp[3] = {(x + 0.5) * lvs[0], (y + 0.5) * lvs[1], (z + 0.5) * lvs[2]};
  double swc = 0, sw = 0;
  for (int i = -3; i <= 3; ++i)
    for (int j = -3; j <= 3; ++j)
      for (int k = -3; k <= 3; ++k) {
        double f[3] = {i * Si, j * Si, k * Si};
        double wk = (Si * Si * Si) * std::exp(-(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) / (2.0 * Si * Si));
        double tt[3] = {(gp[0] + f[0]) / G[0], (gp[1] + f[1]) / G[1], (gp[2] + f[2]) / G[2]};
        double ck = 0;
        if (!(tt[0] < 0 || tt[1] < 0 || tt[2] < 0 || tt[0] > 1 || tt[1] > 1 || tt[2] > 1))
          ck = tex3d_grid<double, double>(prev, px, py, pz, V3<double>(tt[0], tt[1], tt[2]));
        swc += wk * ck;
        sw += wk;
      }
  return swc / sw;
}
double orc_gauss_ext(void* h, double x, double y, double z, double lod) {
  return gauss_ext<double>(*static_cast<Pyramid*>(h), V3<double>(x, y, z), lod);
}
void orc_cone_schedule(int n, double t0, double tmax, double half_deg, double r_sigma, double sigma0, int levels,
                       double* out) {
  std::vector<ConeStep> s = cone_schedule(n, t0, tmax, half_deg, r_sigma, sigma0, levels);
  for (int j = 0; j < n; ++j) { out[3 * j] = s[j].tm; out[3 * j + 1] = s[j].dt; out[3 * j + 2] = s[j].lod; }
}
void orc_render_cone_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                         void* tf, void* pyr, const ConeC* c, const float* cam, int W, int H, double step, double et,
                         double* out, double* margin, int nthreads, int stride) {
  threads(nthreads);
  render_cone<double>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), *static_cast<Pyramid*>(pyr),
                      cone_from(c), mkcam(cam), W, H, step, et, out, margin, stride);
}
void orc_render_cone_f32(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                         void* tf, void* pyr, const ConeC* c, const float* cam, int W, int H, float step, float et,
                         float* out, int nthreads, int stride) {
  threads(nthreads);
  render_cone<float>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), *static_cast<Pyramid*>(pyr),
                     cone_from(c), mkcam(cam), W, H, step, et, out, nullptr, stride);
}
// Exact occlusion / shadow and the surrogate's features at n samples: pos[3n] (texture
// space), dir[3n] (camera ray directions), light[3]; outputs occ[n], sdw[n], feat[16n].
void orc_cone_points(void* pyr, const ConeC* c, int n, const double* pos, const double* dir, double* occ,
                     double* sdw, double* feat, int nthreads) {
  threads(nthreads);
  const Pyramid& P = *static_cast<Pyramid*>(pyr);
  ConeParams cp = cone_from(c);
  ConeTracer<double> ct(P, cp);
  V3<double> L = normalize(V3<double>(cp.light_fwd[0], cp.light_fwd[1], cp.light_fwd[2]));
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    V3<double> p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    V3<double> d = normalize(V3<double>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
    V3<double> k, up, right;
    occ_frame(d, k, up, right);
    if (occ) occ[i] = ct.occlusion(p, k, up, right);
    if (sdw) sdw[i] = ct.shadow(p);
    if (feat) nn_features<double>(P, cp, p, k, L, feat + size_t(i) * kNFeat);
  }
}
// ---------------------------------------------------------------- directional occlusion shading
// Cone spec after the reference's clamps (the caller applies them): half_deg, packing,
// ui_weight, initial_step, covered, d_sigma, r_sigma.
struct DosConeC {
  double half_deg;
  int max_packing;
  double ui_weight, initial_step, covered, d_sigma, r_sigma;
};
struct DosC {
  int apply_occ, apply_sdw, shadow_type;
  double spot_cos, ka, kd;
  DosConeC occ, sdw;
  double lpos[3], lfwd[3], lup[3], lright[3];
};
static DosConeSpec spec_from(const DosConeC& c) {
  DosConeSpec s;
  s.half_deg = c.half_deg; s.max_packing = c.max_packing; s.ui_weight = c.ui_weight;
  s.initial_step = c.initial_step; s.covered = c.covered; s.d_sigma = c.d_sigma; s.r_sigma = c.r_sigma;
  return s;
}
static DosParamsO dos_from(const DosC* c) {
  DosParamsO p;
  p.apply_occ = c->apply_occ; p.apply_sdw = c->apply_sdw; p.shadow_type = c->shadow_type;
  p.spot_cos = c->spot_cos; p.ka = c->ka; p.kd = c->kd;
  p.occ = spec_from(c->occ);
  p.sdw = spec_from(c->sdw);
  for (int i = 0; i < 3; ++i) {
    p.lpos[i] = c->lpos[i]; p.lfwd[i] = c->lfwd[i]; p.lup[i] = c->lup[i]; p.lright[i] = c->lright[i];
  }
  return p;
}
// Section table: returns the row count (or -1); rows[8 i ..] = interval, level, d_integral,
// amplitude, pos, radius, sigma, n_gaussians (up to max_rows); counts[3], axes[30], adj7[1].
int orc_dos_sections(const DosConeC* c, double sigma0, double* rows, int max_rows, int* counts, double* axes,
                     double* adj7) {
  DosTableO T;
  if (!dos_build(spec_from(*c), sigma0, T)) return -1;
  for (int i = 0; i < int(T.s.size()) && i < max_rows; ++i) {
    const DosSection& s = T.s[i];
    double r[8] = {s.interval, s.level, s.d_integral, s.amplitude, s.pos, s.radius, s.sigma, double(s.ng)};
    for (int k = 0; k < 8; ++k) rows[8 * i + k] = r[k];
  }
  for (int k = 0; k < 3; ++k) counts[k] = T.n[k];
  for (int j = 0; j < 10; ++j)
    for (int k = 0; k < 3; ++k) axes[3 * j + k] = T.axes[j][k];
  *adj7 = T.adj7;
  return int(T.s.size());
}
// One cone through a table at texture-space point pos with frame (k, u, v).
double orc_dos_cone(void* pyr, const DosConeC* c, double sigma0, const double* pos, const double* k, const double* u,
                    const double* v) {
  DosTableO T;
  DosConeSpec sp = spec_from(*c);
  if (!dos_build(sp, sigma0, T)) return -1.0;
  return dos_cone<double>(*static_cast<Pyramid*>(pyr), T, sp, V3<double>(pos[0], pos[1], pos[2]),
                          V3<double>(k[0], k[1], k[2]), V3<double>(u[0], u[1], u[2]), V3<double>(v[0], v[1], v[2]));
}
void orc_render_dos_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                        void* tf, void* pyr, const DosC* c, double sigma0, const float* cam, int W, int H,
                        double step, double et, double* out, double* margin, int nthreads, int stride) {
  threads(nthreads);
  render_dos<double>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), *static_cast<Pyramid*>(pyr),
                     dos_from(c), sigma0, mkcam(cam), W, H, step, et, out, margin, stride);
}
void orc_render_dos_f32(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                        void* tf, void* pyr, const DosC* c, double sigma0, const float* cam, int W, int H, float step,
                        float et, float* out, int nthreads, int stride) {
  threads(nthreads);
  render_dos<float>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), *static_cast<Pyramid*>(pyr),
                    dos_from(c), sigma0, mkcam(cam), W, H, step, et, out, nullptr, stride);
}

// ---------------------------------------------------------------- cone ground truth
struct GtC {
  int apply_occ, apply_sdw, shadow_type, n_occ, n_sdw;
  double occ_aperture_deg, sdw_aperture_deg, occ_dist, sdw_dist, gap, rstep, ka, kd;
  double lpos[3], lfwd[3], lup[3], lright[3];
};
static GtParamsO gt_from(const GtC* c) {
  GtParamsO q;
  q.apply_occ = c->apply_occ; q.apply_sdw = c->apply_sdw; q.shadow_type = c->shadow_type;
  std::default_random_engine gen;
  gt_draw(gen, c->n_occ, float(c->occ_aperture_deg), q.occ_rays);
  gt_draw(gen, c->n_sdw, float(c->sdw_aperture_deg), q.sdw_rays);
  q.occ_dist = c->occ_dist; q.sdw_dist = c->sdw_dist; q.gap = c->gap; q.rstep = c->rstep;
  q.ka = c->ka; q.kd = c->kd;
  for (int i = 0; i < 3; ++i) {
    q.lpos[i] = c->lpos[i]; q.lfwd[i] = c->lfwd[i]; q.lup[i] = c->lup[i]; q.lright[i] = c->lright[i];
  }
  return q;
}
void orc_gt_rays(const GtC* c, double* occ, double* sdw) {
  GtParamsO q = gt_from(c);
  std::copy(q.occ_rays.begin(), q.occ_rays.end(), occ);
  std::copy(q.sdw_rays.begin(), q.sdw_rays.end(), sdw);
}
// Transmittance of one ray from texture-space point pos along unit w.
double orc_gt_ray(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz, void* tf,
                  const double* pos, const double* w, double dist, double gap, double rstep) {
  Volume v = mkvol(nx, ny, nz, dtype, data, sx, sy, sz);
  V3<double> G(nx * sx, ny * sy, nz * sz);
  return gt_ray<double>(v, *static_cast<TF*>(tf), G, V3<double>(pos[0], pos[1], pos[2]), V3<double>(w[0], w[1], w[2]),
                        dist, gap, rstep, nullptr);
}
void orc_render_gt_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz, void* tf,
                       const GtC* c, const float* cam, int W, int H, double step, double et, double* out,
                       double* margin, int nthreads, int stride) {
  threads(nthreads);
  render_gt<double>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), gt_from(c), mkcam(cam), W, H,
                    step, et, out, margin, stride);
}
void orc_render_gt_f32(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz, void* tf,
                       const GtC* c, const float* cam, int W, int H, float step, float et, float* out, int nthreads,
                       int stride) {
  threads(nthreads);
  render_gt<float>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), gt_from(c), mkcam(cam), W, H,
                   step, et, out, nullptr, stride);
}

// Light caches (2 doubles per cell, (occlusion, shadow), x fastest) and the march that
// reads them (cache given as 2 doubles per cell).
void orc_lc_ebs(int nx, int ny, int nz, double sx, double sy, double sz, const double* sat, const EbsC* prm, int rx,
                int ry, int rz, double* out, int nthreads) {
  threads(nthreads);
  LightCacheO lc;
  lc_ebs(mkvol(nx, ny, nz, 0, nullptr, sx, sy, sz), sat, ebs_from(prm), rx, ry, rz, lc);
  for (size_t i = 0; i < lc.occ.size(); ++i) { out[2 * i] = lc.occ[i]; out[2 * i + 1] = lc.sdw[i]; }
}
int orc_lc_dos(int nx, int ny, int nz, double sx, double sy, double sz, void* pyr, const DosC* c, double sigma0,
               const float* cam, int W, int H, int rx, int ry, int rz, double* out, int nthreads) {
  threads(nthreads);
  LightCacheO lc;
  bool ok = false;
  lc_dos(mkvol(nx, ny, nz, 0, nullptr, sx, sy, sz), *static_cast<Pyramid*>(pyr), dos_from(c), sigma0, mkcam(cam), W, H,
         rx, ry, rz, lc, &ok);
  if (!ok) return -1;
  for (size_t i = 0; i < lc.occ.size(); ++i) { out[2 * i] = lc.occ[i]; out[2 * i + 1] = lc.sdw[i]; }
  return 0;
}
static LightCacheO lc_from(const double* cache, int rx, int ry, int rz) {
  LightCacheO lc;
  lc.rx = rx; lc.ry = ry; lc.rz = rz;
  const size_t n = size_t(rx) * ry * rz;
  lc.occ.resize(n);
  lc.sdw.resize(n);
  for (size_t i = 0; i < n; ++i) { lc.occ[i] = cache[2 * i]; lc.sdw[i] = cache[2 * i + 1]; }
  return lc;
}
void orc_render_lc_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz, void* tf,
                       const double* cache, int rx, int ry, int rz, int apply_occ, int apply_sdw, double ka, double kd,
                       const float* cam, int W, int H, double step, double et, double* out, double* margin,
                       int nthreads, int stride) {
  threads(nthreads);
  render_lc<double>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), lc_from(cache, rx, ry, rz),
                    apply_occ, apply_sdw, ka, kd, mkcam(cam), W, H, step, et, out, margin, stride);
}
void orc_render_lc_f32(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz, void* tf,
                       const double* cache, int rx, int ry, int rz, int apply_occ, int apply_sdw, double ka, double kd,
                       const float* cam, int W, int H, float step, float et, float* out, int nthreads, int stride) {
  threads(nthreads);
  render_lc<float>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), lc_from(cache, rx, ry, rz),
                   apply_occ, apply_sdw, ka, kd, mkcam(cam), W, H, step, et, out, nullptr, stride);
}

// Gradients (3 doubles per voxel, x fastest) and the Blinn-Phong renders: the gradient
// of `type` (0 Sobel-Feldman, 1 finite differences) is built from the volume itself.
struct PhongC {
  int type;
  double ka, kd, ks, shininess;
  double spec[3], lpos[3];
};
static void grad_of(const Volume& v, int type, GradientO& g) {
  if (type == 0) gradient_sobel(v, g);
  else gradient_fd(v, g);
}
static PhongO phong_from(const PhongC* c, const GradientO* g) {
  PhongO p;
  p.g = g;
  p.ka = c->ka; p.kd = c->kd; p.ks = c->ks; p.shininess = c->shininess;
  for (int i = 0; i < 3; ++i) { p.spec[i] = c->spec[i]; p.lpos[i] = c->lpos[i]; }
  return p;
}
void orc_gradient(int nx, int ny, int nz, int dtype, const void* data, int type, double* out, int nthreads) {
  threads(nthreads);
  GradientO g;
  grad_of(mkvol(nx, ny, nz, dtype, data, 1, 1, 1), type, g);
  for (size_t i = 0; i < g.gx.size(); ++i) { out[3 * i] = g.gx[i]; out[3 * i + 1] = g.gy[i]; out[3 * i + 2] = g.gz[i]; }
}
void orc_render_rc1p_phong_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                               void* tf, const PhongC* ph, const float* cam, int W, int H, double step, double et,
                               double* out, double* margin, int nthreads, int stride) {
  threads(nthreads);
  const Volume v = mkvol(nx, ny, nz, dtype, data, sx, sy, sz);
  GradientO g;
  grad_of(v, ph->type, g);
  render_rc1p_phong<double>(v, *static_cast<TF*>(tf), phong_from(ph, &g), mkcam(cam), W, H, step, et, out, margin,
                            stride);
}
void orc_render_ebs_phong_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                              void* tf, const double* sat, const EbsC* prm, const PhongC* ph, const float* cam, int W,
                              int H, double step, double et, double* out, double* margin, int nthreads, int stride) {
  threads(nthreads);
  const Volume v = mkvol(nx, ny, nz, dtype, data, sx, sy, sz);
  GradientO g;
  grad_of(v, ph->type, g);
  const PhongO P = phong_from(ph, &g);
  render_ebs<double, double>(v, *static_cast<TF*>(tf), sat, ebs_from(prm), mkcam(cam), W, H, step, et, out, margin,
                             stride, &P);
}
void orc_render_dos_phong_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                              void* tf, void* pyr, const DosC* c, double sigma0, const PhongC* ph, const float* cam,
                              int W, int H, double step, double et, double* out, double* margin, int nthreads,
                              int stride) {
  threads(nthreads);
  const Volume v = mkvol(nx, ny, nz, dtype, data, sx, sy, sz);
  GradientO g;
  grad_of(v, ph->type, g);
  const PhongO P = phong_from(ph, &g);
  render_dos<double>(v, *static_cast<TF*>(tf), *static_cast<Pyramid*>(pyr), dos_from(c), sigma0, mkcam(cam), W, H, step,
                     et, out, margin, stride, &P);
}
void orc_render_lc_phong_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz,
                             void* tf, const double* cache, int rx, int ry, int rz, int apply_occ, int apply_sdw,
                             double ka, double kd, const PhongC* ph, const float* cam, int W, int H, double step,
                             double et, double* out, double* margin, int nthreads, int stride) {
  threads(nthreads);
  const Volume v = mkvol(nx, ny, nz, dtype, data, sx, sy, sz);
  GradientO g;
  grad_of(v, ph->type, g);
  const PhongO P = phong_from(ph, &g);
  render_lc<double>(v, *static_cast<TF*>(tf), lc_from(cache, rx, ry, rz), apply_occ, apply_sdw, ka, kd, mkcam(cam), W,
                    H, step, et, out, margin, stride, &P);
}

// Multi-scale output: src (sh x sw x 4) resampled into dst (dh x dw x 4).
void orc_resample(const double* src, int sw, int sh, double* dst, int dw, int dh, int mode, int kernel) {
  Frame s, d;
  s.w = sw; s.h = sh;
  s.p.assign(src, src + size_t(sw) * sh * 4);
  d.w = dw; d.h = dh;
  resample(s, d, mode, kernel);
  std::copy(d.p.begin(), d.p.end(), dst);
}

int orc_mlp_nparams() { return kNParams; }
void orc_mlp_forward(const double* prm, int n, const double* X, double* Y) {
  for (int i = 0; i < n; ++i) mlp_forward(prm, X + size_t(i) * kNFeat, Y + size_t(i) * kOut);
}
double orc_mlp_loss_grad(const double* prm, const double* X, const double* T, int n, double* grad) {
  return mlp_loss_grad(prm, X, T, n, grad);
}
void orc_mlp_forward_bf16(const double* prm, int n, const double* X, double* Y) {
  for (int i = 0; i < n; ++i) mlp_forward_bf16(prm, X + size_t(i) * kNFeat, Y + size_t(i) * kOut);
}
// The surrogate-shaded frame (render_nn): prm[kNParams], lo/hi[16] (the range outside
// which the exact tracer shades), bf16 selects the network's rounding points;
// counts[2] receives (shaded samples, exact-tracer samples).
void orc_render_nn_f64(int nx, int ny, int nz, int dtype, const void* data, double sx, double sy, double sz, void* tf,
                       void* pyr, const ConeC* c, const double* prm, const double* lo, const double* hi, int bf16,
                       const float* cam, int W, int H, double step, double et, double* out, double* margin,
                       int nthreads, int stride, long long* counts) {
  threads(nthreads);
  NnNet net{prm, lo, hi, bf16};
  render_nn<double>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), *static_cast<Pyramid*>(pyr),
                    cone_from(c), net, mkcam(cam), W, H, step, et, out, margin, stride, counts);
}

}  // extern "C"
