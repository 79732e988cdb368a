// CPU ORACLE — C entry points for the parity tests (test infrastructure only).
// Built by oracle/Makefile into oracle/_build/liboracle.so; loaded only by tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg.
#include "oracle.hpp"
#include "oracle_cone.hpp"
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
                        float et, float* out, int nthreads) {
  threads(nthreads);
  render_ebs<float, float>(mkvol(nx, ny, nz, dtype, data, sx, sy, sz), *static_cast<TF*>(tf), sat, ebs_from(prm),
                           mkcam(cam), W, H, step, et, out, nullptr);
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

}  // extern "C"
