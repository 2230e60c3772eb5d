// Runs every entry point of the CPU oracle (oracle/oracle.cpp, compiled into
// this program) on small seeded inputs, for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer: any out-of-bounds access, use after free, leak,
// signed overflow or other undefined behaviour stops the program with a
// report. It also checks a few results that hold exactly (the loss gradient
// the backward receives, the PLY round trip), so a silently broken path fails
// too. Exit status 0: clean.
#include "../../oracle/oracle.cpp"

#include <cstdio>
#include <cstdlib>
#include <random>

namespace {

int fails = 0;
void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++fails;
  }
}

std::vector<float> scene(size_t n, uint32_t seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<float> N(0.0f, 1.0f);
  std::vector<float> g(n * 59, 0.0f);
  for (size_t i = 0; i < n; ++i) {
    float* r = &g[i * 59];
    for (int k = 0; k < 3; ++k) r[k] = 0.4f * N(rng);
    r[2] += 2.5f;
    r[3] = N(rng);
    float q[4], s = 0.0f;
    for (int k = 0; k < 4; ++k) q[k] = N(rng), s += q[k] * q[k];
    for (int k = 0; k < 4; ++k) r[4 + k] = q[k] / std::sqrt(s);
    for (int k = 0; k < 3; ++k) r[8 + k] = -3.0f + 0.5f * N(rng);
    for (int k = 0; k < 48; ++k) r[11 + k] = 0.05f * N(rng);
    r[11] += 0.3f * N(rng);
  }
  return g;
}

}  // namespace

int main() {
  orc_set_threads(2);
  const size_t n = 300;
  std::vector<float> g = scene(n, 1), tg = scene(n, 2);
  Camera cam{};
  cam.width = 72;
  cam.height = 40;
  cam.fx = cam.fy = 70.0f;
  cam.cx = 36.0f;
  cam.cy = 20.0f;
  for (int k = 0; k < 4; ++k) cam.c2w[k * 5] = 1.0f;
  const size_t px = (size_t)cam.width * cam.height;
  const float bg[3] = {0.1f, 0.2f, 0.3f};
  int err = 0;

  // target, render, backward, loss, their accessors
  void* ht = orc_run_f32(tg.data(), n, &cam, nullptr, bg, 0.2f, 0, &err);
  check(ht && !err, "target render");
  std::vector<float> target(px * 4, 1.0f), rgb(px * 3), ft(px);
  std::vector<int32_t> last(px);
  orc_get_image(ht, 0, rgb.data(), ft.data(), last.data());
  for (size_t p = 0; p < px; ++p)
    for (int c = 0; c < 3; ++c) target[p * 4 + c] = rgb[p * 3 + c];
  orc_free(ht, 0);
  void* h = orc_run_f32(g.data(), n, &cam, target.data(), bg, 0.2f, 1, &err);
  check(h && !err, "render_and_backward");
  const size_t ns = orc_nsplats(h, 0), np = orc_npairs(h, 0);
  std::vector<uint64_t> gi(ns + 1), off((cam.width + 15) / 16 * ((cam.height + 15) / 16) + 1);
  std::vector<float> fields((ns + 1) * 22);
  std::vector<int32_t> rects((ns + 1) * 4);
  std::vector<uint32_t> gis(np + 1);
  orc_get_splats(h, 0, gi.data(), fields.data(), rects.data());
  orc_get_tiles(h, 0, off.data(), gis.data());
  check(off.back() == np, "tile offsets end at the pair count");
  orc_get_image(h, 0, rgb.data(), ft.data(), last.data());
  std::vector<float> dl(px * 3), grads(n * 60), grads2(n * 60);
  orc_get_dl(h, 0, dl.data());
  orc_get_grads(h, 0, grads.data());
  orc_backward_with_dl_f32(h, g.data(), n, bg, dl.data(), grads2.data());
  check(grads == grads2, "backward with its own loss gradient");
  const double loss = orc_get_loss(h, 0), exact = orc_get_loss_exact(h, 0);
  check(std::fabs(loss - exact) <= 1e-5 * std::fabs(exact), "loss sums");
  (void)orc_pairs_fwd(h, 0);
  orc_free(h, 0);
  std::vector<double> g64(g.begin(), g.end()), t64(target.begin(), target.end());
  const double bg64[3] = {0.1, 0.2, 0.3};
  void* h64 = orc_run_f64(g64.data(), n, &cam, t64.data(), bg64, 0.2, 1, &err);
  check(h64 && !err, "f64 render_and_backward");
  orc_free(h64, 1);
  std::vector<float> rgba(px * 4), dl2(px * 3);
  check(orc_rasterize_rgba(g.data(), n, &cam, bg, rgba.data()) == 0, "rasterize");
  double ex = 0.0;
  (void)orc_loss_grad_f32_exact(rgb.data(), target.data(), cam.width, cam.height, 0.2f, dl2.data(), &ex);
  std::vector<double> rgb64(rgb.begin(), rgb.end()), dl64(px * 3);
  (void)orc_loss_grad_f64(rgb64.data(), t64.data(), cam.width, cam.height, 0.2, dl64.data());

  // optimizer: Adam, opacity reset, densify/prune with growth
  orc_train_cfg cfg{1.6e-4f, 1e-3f, 5e-3f, 5e-2f, 5e-3f, 2.5e-4f, 0.9f, 0.999f, 1e-8f, 2e-4f, 0.005f,
                    500, 15000, 3000, 2 * n, 30000};
  const size_t cap = 3 * n;
  std::vector<float> gb(cap * 59, 0.0f), m1(cap * 59, 0.0f), m2(cap * 59, 0.0f), acc(cap, 0.0f);
  std::vector<uint32_t> den(cap, 0u);
  std::copy(g.begin(), g.end(), gb.begin());
  uint64_t t = 0;
  for (int it = 0; it < 3; ++it) orc_adam_step(gb.data(), grads.data(), n, m1.data(), m2.data(), acc.data(),
                                               den.data(), &t, &cfg);
  check(t == 3, "adam step count");
  for (size_t i = 0; i < n; ++i) acc[i] = 1e-3f * (float)(i % 7), den[i] = 1;
  orc_reset_opacities(gb.data(), n);
  size_t nn = n, rows = n;
  const float extent = orc_scene_extent(gb.data(), n);
  check(orc_densify_and_prune(gb.data(), &nn, cap, m1.data(), m2.data(), acc.data(), den.data(), &rows, &cfg,
                              extent) == 0,
        "densify_and_prune");
  check(nn <= cap && rows <= cap, "densify stays in capacity");

  // model files: PLY round trip, .splat
  const char* ply = "/tmp/oracle_sanitize.ply";
  check(orc_save_ply(g.data(), n, ply) == 0, "save_ply");
  std::vector<float> back(n * 59);
  size_t nb = 0;
  check(orc_load_ply(ply, back.data(), n, &nb) == 0 && nb == n && back == g, "PLY round trip");
  check(orc_save_splat(g.data(), n, "/tmp/oracle_sanitize.splat") == 0, "save_splat");
  std::remove(ply);
  std::remove("/tmp/oracle_sanitize.splat");

  // colour, resize, preview, SfM initialisation
  std::vector<uint8_t> img(37 * 23 * 4), small(20 * 11 * 4), bytes(px * 4);
  for (size_t i = 0; i < img.size(); ++i) img[i] = (uint8_t)(i * 37 + 11);
  check(orc_resize_rgba8(img.data(), 37, 23, small.data(), 20, 11) == 0, "resize down");
  std::vector<uint8_t> big(50 * 40 * 4);
  check(orc_resize_rgba8(img.data(), 37, 23, big.data(), 50, 40) == 0, "resize up");
  std::vector<float> lin(img.size());
  orc_rgba8_to_linear(img.data(), img.size() / 4, lin.data());
  orc_to_u8_srgb(rgba.data(), px, bytes.data());
  std::vector<float> pv(n * 10), prgb(px * 3);
  check(orc_preview_project(g.data(), n, &cam, pv.data()) == 0, "preview project");
  check(orc_preview_render(g.data(), n, &cam, bg, prgb.data()) == 0, "preview render");
  std::vector<double> xyz(n * 3);
  std::vector<uint8_t> col(n * 3);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) xyz[i * 3 + k] = (double)g[i * 59 + k], col[i * 3 + k] = (uint8_t)(i * 3 + k);
  std::vector<float> sfm(n * 59), scales(n);
  check(orc_gaussians_from_sfm(xyz.data(), col.data(), n, sfm.data(), scales.data()) == 0, "gaussians_from_sfm");

  // math helpers
  float q[4] = {0.3f, -0.2f, 0.5f, 0.7f}, o9[9], o36[36], o16[16], o48[48], o3[3], d3[3] = {0.2f, -0.3f, 0.9f};
  orc_to_rotation_matrix(q, o9);
  orc_rotation_jacobian(q, o36);
  orc_sh_basis(d3, o16);
  orc_sh_basis_grad(d3, o48);
  orc_eval_sh_deg3(&g[11], d3, o3);
  float view[16];
  check(orc_view_matrix(&cam, view) == 0, "view matrix");
  std::vector<float> xs(64), es(64);
  for (int i = 0; i < 64; ++i) xs[i] = -100.0f + 3.3f * (float)i;
  orc_expf_batch(xs.data(), es.data(), xs.size());
  (void)orc_powi(0.999f, 3000);
  if (fails) return 1;
  std::printf("oracle sanitizer run clean: %zu splats, %zu pairs, loss %.6f\n", ns, np, loss);
  return 0;
}
