// Runs every entry point of the oracle once on small seeded scenes with edge
// inputs (off-screen and culled Gaussians, inverted and NaN boxes, empty
// tiles, ties), for a build with AddressSanitizer and UBSan: any out-of-
// bounds access, use of freed memory or undefined operation aborts the run.
// Test infrastructure, like the oracle itself (run by
// tests/test_oracle_sanitize.py through `make -C oracle sanitize`).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" {
uint32_t orc_gauss_project_f32(uint32_t, const float*, const float*, const float*, const float*,
                               uint32_t, uint32_t, float*, float*, float*);
uint32_t orc_gauss_project_f64(uint32_t, const double*, const double*, const double*,
                               const double*, uint32_t, uint32_t, double*, double*, double*);
uint64_t orc_tile2pnt(uint32_t, const float*, const float*, uint32_t, uint32_t, uint32_t,
                      uint32_t*, uint32_t*, uint64_t*);
void orc_circle_boxes(uint32_t, const float*, const float*, float*);
void orc_gauss_rasterize_f32(uint32_t, uint32_t, const float*, const float*, const float*,
                             const float*, const float*, uint32_t, const uint32_t*,
                             const uint32_t*, float*, float*, uint32_t*, float*, const uint32_t*);
void orc_gauss_rasterize_exp_ulp_f32(uint32_t, uint32_t, const float*, const float*, const float*,
                                     const float*, const float*, uint32_t, const uint32_t*,
                                     const uint32_t*, float*, float*, uint32_t*, float*, int);
void orc_gauss_rasterize_f64(uint32_t, uint32_t, const double*, const double*, const float*,
                             const double*, const double*, uint32_t, const uint32_t*,
                             const uint32_t*, double*, double*, uint32_t*, double*,
                             const uint32_t*);
void orc_gauss_rasterize_bwd_f32(uint32_t, uint32_t, const float*, const float*, const float*,
                                 const float*, const float*, uint32_t, const uint32_t*,
                                 const uint32_t*, const uint32_t*, const float*, const float*,
                                 const float*, float*, float*, float*, float*);
void orc_gauss_rasterize_bwd_mt_f32(uint32_t, uint32_t, uint32_t, const float*, const float*,
                                    const float*, const float*, const float*, uint32_t,
                                    const uint32_t*, const uint32_t*, const uint32_t*,
                                    const float*, const float*, const float*, float*, float*,
                                    float*, float*);
void orc_gauss_project_bwd_f32(uint32_t, const float*, const float*, const float*, const float*,
                               uint32_t, uint32_t, const float*, const float*, float*, float*,
                               float*, int);
void orc_gauss_render_sort_depth(uint32_t, const float*, const float*, const float*,
                                 const float*, const float*, uint32_t, uint32_t, float*);
void orc_circle_project(uint32_t, const float*, float, const float*, uint32_t, uint32_t, float*,
                        float*);
void orc_circle_rasterize(uint32_t, uint32_t, const float*, const float*, const float*, uint32_t,
                          const uint32_t*, const uint32_t*, float*);
void orc_zbuffer(uint32_t, const float*, const float*, const float*, uint32_t, uint32_t,
                 uint64_t*, float*);
void orc_render_pix_sort_depth(uint32_t, const float*, const float*, uint32_t, uint32_t, float*);
void orc_adam(uint32_t, float*, const float*, float*, float*, float, float, float, float,
              uint32_t);
void orc_sh_to_rgb_f32(uint32_t, uint32_t, uint32_t, const float*, const float*, const float*,
                       float*);
void orc_sh_to_rgb_bwd_f32(uint32_t, uint32_t, uint32_t, const float*, const float*,
                           const float*, const float*, float*, float*, int);
int orc_set_threads(int);
}

namespace {

std::mt19937 rng(20261204u);
float uni(float a, float b) { return std::uniform_real_distribution<float>(a, b)(rng); }
float gauss() { return std::normal_distribution<float>(0.f, 1.f)(rng); }

// Column-major world -> ndc of a camera on the +z axis at `dist`, looking at
// the origin (fov 60 degrees, aspect w / h, near 0.01, far 100).
std::vector<float> camera(uint32_t w, uint32_t h, float dist) {
  const float f = 1.f / std::tan(3.14159265f / 6.f), a = float(w) / float(h);
  const float n = 0.01f, fa = 100.f;
  std::vector<float> m(16, 0.f);
  m[0] = f / a;
  m[5] = f;
  m[10] = (fa + n) / (n - fa);
  m[11] = -1.f;
  m[14] = 2.f * fa * n / (n - fa) + (fa + n) / (n - fa) * -dist;
  m[15] = dist;
  return m;
}

int fail(const char* what) {
  std::fprintf(stderr, "sanitize_check: %s\n", what);
  return 1;
}

}  // namespace

int main() {
  const uint32_t w = 83, h = 61, n = 700;
  std::vector<float> xyz(3 * n), quat(4 * n), scale(3 * n), op(n), rgb(3 * n);
  for (uint32_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      xyz[3 * i + c] = uni(-1.f, 1.f);
      scale[3 * i + c] = std::exp(-3.f + 0.5f * gauss());
      rgb[3 * i + c] = uni(0.f, 1.f);
    }
    float qn = 0.f;
    for (int c = 0; c < 4; ++c) {
      quat[4 * i + c] = gauss();
      qn += quat[4 * i + c] * quat[4 * i + c];
    }
    for (int c = 0; c < 4; ++c) quat[4 * i + c] /= std::sqrt(qn);
    op[i] = 1.f / (1.f + std::exp(-gauss()));
  }
  // edge Gaussians: behind the camera (culled), far off-screen, huge, tiny,
  // and two at the same depth
  xyz[0] = 0.f, xyz[1] = 0.f, xyz[2] = 5.f;
  xyz[3] = 40.f, xyz[4] = -30.f;
  scale[6] = scale[7] = scale[8] = 3.f;
  scale[9] = scale[10] = scale[11] = 1e-6f;
  xyz[14] = xyz[17];
  const std::vector<float> m = camera(w, h, 3.f);
  std::vector<float> pxd(3 * n), conic(3 * n), aabb(4 * n);
  orc_gauss_project_f32(n, xyz.data(), quat.data(), scale.data(), m.data(), w, h, pxd.data(),
                        conic.data(), aabb.data());
  {
    std::vector<double> xd(xyz.begin(), xyz.end()), qd(quat.begin(), quat.end()),
        sd(scale.begin(), scale.end()), md(m.begin(), m.end());
    std::vector<double> p64(3 * n), c64(3 * n), b64(4 * n);
    orc_gauss_project_f64(n, xd.data(), qd.data(), sd.data(), md.data(), w, h, p64.data(),
                          c64.data(), b64.data());
  }
  // boxes of every kind for the binning alone
  std::vector<float> boxes(aabb);
  const float odd[][4] = {{40, 30, 35, 20}, {40, 30, 20, 20}, {NAN, 0, 10, 10},
                          {-1e30f, -1e30f, 1e30f, 1e30f}, {INFINITY, INFINITY, -INFINITY, -INFINITY}};
  for (int k = 0; k < 5; ++k)
    for (int c = 0; c < 4; ++c) boxes[4 * (20 + k) + c] = odd[k][c];
  for (uint32_t ts : {8u, 16u, 32u}) {
    const uint32_t nt = ((w + ts - 1) / ts) * ((h + ts - 1) / ts);
    std::vector<uint32_t> t2i(nt + 1);
    const uint64_t K = orc_tile2pnt(n, boxes.data(), pxd.data(), w, h, ts, t2i.data(), nullptr,
                                    nullptr);
    std::vector<uint32_t> i2p(K + 1);
    std::vector<uint64_t> keys(K + 1);
    if (orc_tile2pnt(n, boxes.data(), pxd.data(), w, h, ts, t2i.data(), i2p.data(),
                     keys.data()) != K || t2i[nt] != K)
      return fail("tile2pnt count mismatch");
  }
  const uint32_t ts = 16, nt = ((w + ts - 1) / ts) * ((h + ts - 1) / ts);
  std::vector<uint32_t> t2i(nt + 1);
  const uint64_t K = orc_tile2pnt(n, aabb.data(), pxd.data(), w, h, ts, t2i.data(), nullptr,
                                  nullptr);
  std::vector<uint32_t> i2p(K + 1);
  orc_tile2pnt(n, aabb.data(), pxd.data(), w, h, ts, t2i.data(), i2p.data(), nullptr);
  const size_t P = size_t(w) * h;
  std::vector<float> img(3 * P), alpha(P), tfinal(P), target(3 * P);
  std::vector<uint32_t> nwalk(P);
  orc_gauss_rasterize_f32(w, h, pxd.data(), conic.data(), aabb.data(), op.data(), rgb.data(), ts,
                          t2i.data(), i2p.data(), img.data(), alpha.data(), nwalk.data(),
                          tfinal.data(), nullptr);
  orc_gauss_rasterize_exp_ulp_f32(w, h, pxd.data(), conic.data(), aabb.data(), op.data(),
                                  rgb.data(), ts, t2i.data(), i2p.data(), img.data(),
                                  alpha.data(), nwalk.data(), tfinal.data(), 1);
  orc_gauss_rasterize_f32(w, h, pxd.data(), conic.data(), aabb.data(), op.data(), rgb.data(), ts,
                          t2i.data(), i2p.data(), img.data(), alpha.data(), nwalk.data(),
                          tfinal.data(), nullptr);
  {
    std::vector<double> p64(pxd.begin(), pxd.end()), c64(conic.begin(), conic.end()),
        o64(op.begin(), op.end()), r64(rgb.begin(), rgb.end()), img64(3 * P), a64(P), t64(P);
    std::vector<uint32_t> nw64(P);
    orc_gauss_rasterize_f64(w, h, p64.data(), c64.data(), aabb.data(), o64.data(), r64.data(), ts,
                            t2i.data(), i2p.data(), img64.data(), a64.data(), nw64.data(),
                            t64.data(), nwalk.data());
  }
  for (size_t k = 0; k < 3 * P; ++k) target[k] = 2.f * (img[k] - uni(0.f, 1.f));
  std::vector<float> dpix(2 * n), dcon(3 * n), dop(n), drgb(3 * n), dalpha(P, 0.25f);
  orc_gauss_rasterize_bwd_f32(w, h, pxd.data(), conic.data(), aabb.data(), op.data(), rgb.data(),
                              ts, t2i.data(), i2p.data(), nwalk.data(), tfinal.data(),
                              target.data(), dalpha.data(), dpix.data(), dcon.data(), dop.data(),
                              drgb.data());
  const int nth = orc_set_threads(4);
  std::vector<float> dpix2(2 * n), dcon2(3 * n), dop2(n), drgb2(3 * n);
  if (nth < 1) return fail("set_threads");
  orc_gauss_rasterize_bwd_mt_f32(n, w, h, pxd.data(), conic.data(), aabb.data(),
                                 op.data(), rgb.data(), ts, t2i.data(), i2p.data(), nwalk.data(),
                                 tfinal.data(), target.data(), nullptr, dpix2.data(), dcon2.data(),
                                 dop2.data(), drgb2.data());
  orc_set_threads(1);
  std::vector<float> dxyz(3 * n), dsc(3 * n), dq(4 * n);
  orc_gauss_project_bwd_f32(n, xyz.data(), quat.data(), scale.data(), m.data(), w, h, dpix.data(),
                            dcon.data(), dxyz.data(), dsc.data(), dq.data(), 0);
  orc_gauss_render_sort_depth(n, pxd.data(), conic.data(), aabb.data(), op.data(), rgb.data(), w,
                              h, img.data());
  // circles and points
  std::vector<float> cpxd(3 * n), crad(n), cbox(4 * n);
  orc_circle_project(n, xyz.data(), 0.02f, m.data(), w, h, cpxd.data(), crad.data());
  for (uint32_t i = 0; i < n; ++i) crad[i] = uni(0.f, 4.f);
  crad[5] = -1.f;
  crad[6] = INFINITY;
  cpxd[21] = NAN;
  cpxd[24] = 1e30f;
  orc_circle_boxes(n, cpxd.data(), crad.data(), cbox.data());
  const uint64_t Kc = orc_tile2pnt(n, cbox.data(), cpxd.data(), w, h, ts, t2i.data(), nullptr,
                                   nullptr);
  std::vector<uint32_t> ci2p(Kc + 1);
  orc_tile2pnt(n, cbox.data(), cpxd.data(), w, h, ts, t2i.data(), ci2p.data(), nullptr);
  orc_circle_rasterize(w, h, cpxd.data(), crad.data(), rgb.data(), ts, t2i.data(), ci2p.data(),
                       img.data());
  std::vector<uint64_t> zkeys(P);
  orc_zbuffer(n, cpxd.data(), crad.data(), rgb.data(), w, h, zkeys.data(), img.data());
  orc_zbuffer(n, cpxd.data(), nullptr, rgb.data(), w, h, zkeys.data(), img.data());
  orc_render_pix_sort_depth(n, cpxd.data(), rgb.data(), w, h, img.data());
  // SH colour at every coefficient count, and Adam
  for (uint32_t K2 : {1u, 4u, 9u, 16u}) {
    const uint32_t deg = K2 == 1 ? 0 : K2 == 4 ? 1 : K2 == 9 ? 2 : 3;
    std::vector<float> sh(size_t(n) * K2 * 3), dsh(sh.size()), out(3 * n), dx(3 * n);
    for (float& v : sh) v = 0.3f * gauss();
    const float cam[3] = {0.f, 0.f, 3.f};
    orc_sh_to_rgb_f32(n, deg, K2, cam, xyz.data(), sh.data(), out.data());
    orc_sh_to_rgb_bwd_f32(n, deg, K2, cam, xyz.data(), sh.data(), drgb.data(), dsh.data(),
                          dx.data(), 1);
  }
  std::vector<float> m1(3 * n), m2(3 * n);
  for (uint32_t step = 1; step <= 3; ++step)
    orc_adam(3 * n, xyz.data(), dxyz.data(), m1.data(), m2.data(), 1e-3f, 0.9f, 0.999f, 1e-8f,
             step);
  std::printf("sanitize_check: ok (K = %llu at 16x16 tiles, %llu circle pairs)\n",
              static_cast<unsigned long long>(K), static_cast<unsigned long long>(Kc));
  return 0;
}
