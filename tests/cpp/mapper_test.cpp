// mapper_test.cpp -- checks of the C++ mapper layer (rvk_mapper.h).
//
//   rvk_mapper_test cpu   transfer-function tables, polygons, .2tf files and
//                         error paths (no device needed)
//   rvk_mapper_test gpu   mapper renders against direct C-ABI renders
//
// Each check prints "PASS <name>" or "FAIL <name>: <why>"; the exit status is
// the number of failures (capped at 100).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include <unistd.h>

#include "rvk_mapper.h"
#include "rvk_synth.h"

namespace {

int g_fail = 0;

void report(const std::string& name, bool ok, const std::string& why = "") {
  if (ok) {
    std::printf("PASS %s\n", name.c_str());
  } else {
    std::printf("FAIL %s: %s\n", name.c_str(), why.c_str());
    ++g_fail;
  }
  std::fflush(stdout);
}

std::string fmt(const char* f, double a, double b = 0, double c = 0) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), f, a, b, c);
  return buf;
}

// ---------------------------------------------------------------------------------
void cpu_piecewise() {
  rvk::PiecewiseFunction f;
  f.AddPoint(10, 1.0);
  f.AddPoint(0, 0.0);
  f.AddPoint(20, 0.5);
  f.AddPoint(10, 0.8);  // replaces the value at x = 10
  double lo, hi;
  f.GetRange(lo, hi);
  bool ok = lo == 0 && hi == 20;
  ok &= f.GetValue(-5) == 0.0 && f.GetValue(25) == 0.5;  // clamped ends
  ok &= std::fabs(f.GetValue(5) - 0.4) < 1e-15 && std::fabs(f.GetValue(15) - 0.65) < 1e-15;
  float t[5];
  f.GetTable(0, 20, 5, t);  // samples at 0, 5, 10, 15, 20
  const float want[5] = {0.0f, 0.4f, 0.8f, 0.65f, 0.5f};
  for (int i = 0; i < 5; ++i) ok &= std::fabs(t[i] - want[i]) < 1e-7f;
  report("piecewise_function", ok);

  rvk::ColorTransferFunction c;
  c.AddRGBPoint(0, 0, 0.5, 1);
  c.AddRGBPoint(100, 1, 0.5, 0);
  float rgb[9];
  c.GetTable(0, 100, 3, rgb);
  ok = std::fabs(rgb[3] - 0.5f) < 1e-7f && rgb[4] == 0.5f && std::fabs(rgb[5] - 0.5f) < 1e-7f && rgb[0] == 0 &&
       rgb[8] == 0 && rgb[6] == 1;
  report("color_transfer_function", ok);
}

std::shared_ptr<rvk::FunctionPolygon> rect(float a, float b, float c, float d) {
  auto p = std::make_shared<rvk::FunctionPolygon>();
  p->AddVertex(a, c);
  p->AddVertex(b, c);
  p->AddVertex(b, d);
  p->AddVertex(a, d);
  return p;
}

void cpu_polygons() {
  // Even-odd test of an axis-aligned rectangle: inside iff a < x <= b and
  // c < y <= d (crossings counted on edges whose gradient span is half-open).
  auto r = rect(2, 6, 3, 8);
  bool ok = r->PointInPolygon(4, 5) && !r->PointInPolygon(1, 5) && !r->PointInPolygon(7, 5) &&
            !r->PointInPolygon(4, 2) && !r->PointInPolygon(4, 9);
  ok &= r->PointInPolygon(6, 8) && !r->PointInPolygon(2, 5) && !r->PointInPolygon(4, 3);
  // a triangle (0,0) (10,0) (0,10): inside iff x > 0, y > 0, x + y < 10 away from edges
  rvk::FunctionPolygon tri;
  tri.AddVertex(0, 0);
  tri.AddVertex(10, 0);
  tri.AddVertex(0, 10);
  ok &= tri.PointInPolygon(2, 3) && !tri.PointInPolygon(6, 6) && !tri.PointInPolygon(-1, 1);
  // a concave "U" shape: the notch is outside
  rvk::FunctionPolygon u;
  const float uv[8][2] = {{0, 0}, {9, 0}, {9, 9}, {6, 9}, {6, 3}, {3, 3}, {3, 9}, {0, 9}};
  for (auto& v : uv) u.AddVertex(v[0], v[1]);
  ok &= u.PointInPolygon(1.5f, 6) && u.PointInPolygon(7.5f, 6) && !u.PointInPolygon(4.5f, 6) &&
        u.PointInPolygon(4.5f, 1.5f);
  ok &= r->getMinIntensity() == 2 && r->getMaxIntensity() == 6 && r->getMinGradient() == 3 &&
        r->getMaxGradient() == 8;
  report("point_in_polygon", ok);

  // Linear-axis rasterisation over an 11 x 11 table spanning [0,10] on both
  // axes (sample k sits at k): the polygon fills its bounding-box indices
  // [low, high) that pass the point test, i.e. i in {3,4,5}, g in {4..7}.
  const int n = 11;
  std::vector<float> R(n * n, -1), G(n * n, -1), B(n * n, -1), A(n * n, -1);
  r->SetColour(0.25f, 0.5f, 0.75f);
  r->SetOpacity(0.125f);
  r->PopulatePortionOfTransferTable(n, n, 0, 10, 0, 0, 10, 0, R.data(), G.data(), B.data(), A.data(), 0);
  ok = true;
  for (int g = 0; g < n; ++g)
    for (int i = 0; i < n; ++i) {
      const bool in = i >= 3 && i <= 5 && g >= 4 && g <= 7;
      const int k = i + g * n;
      if (in) ok &= R[k] == 0.25f && G[k] == 0.5f && B[k] == 0.75f && A[k] == 0.125f;
      else ok &= R[k] == -1 && A[k] == -1;
    }
  report("rasterise_linear_rectangle", ok);

  // Log gradient axis (logUsed = 2): row g samples the gradient
  // sqrt(2^v off - off) with v = lowG + g (highG - lowG)/(n - 1); the
  // rectangle then covers exactly the rows whose gradient lies in (c, d]
  // inside its bounding rows.
  const float off = 4.0f, lowG = 0.0f, highG = 8.0f;
  std::vector<float> amb(n * n, -1), dif(n * n, -1), spec(n * n, -1), pw(n * n, -1);
  r->SetAmbient(0.5f);
  r->SetDiffuse(0.25f);
  r->SetSpecular(0.125f);
  r->SetSpecularPower(7.0f);
  r->PopulatePortionOfShadingTable(n, n, 0, 10, 0, lowG, highG, off, amb.data(), dif.data(), spec.data(), pw.data(),
                                   2);
  auto v_of = [&](double gr) { return (std::log(gr * gr + off) - std::log(off)) / std::log(2.0); };
  const int lg = (int)((n - 1) * (v_of(3) - lowG) / (highG - lowG));
  const int hg = (int)((n - 1) * (v_of(8) - lowG) / (highG - lowG));
  ok = true;
  int filled = 0;
  for (int g = 0; g < n; ++g) {
    const double v = g * (highG - lowG) / (double)(n - 1) + lowG;
    const double gr = std::sqrt(std::exp(std::log(2.0) * v + std::log((double)off)) - off);
    for (int i = 0; i < n; ++i) {
      const bool in = g >= lg && g < hg && gr > 3 && gr <= 8 && i >= 3 && i <= 5;
      const int k = i + g * n;
      if (in) {
        ok &= amb[k] == 0.5f && dif[k] == 0.25f && spec[k] == 0.125f && pw[k] == 7.0f;
        ++filled;
      } else {
        ok &= amb[k] == -1;
      }
    }
  }
  report("rasterise_log_gradient_rectangle", ok && filled > 0, fmt("filled %g", filled));

  // Composite function: later objects overwrite earlier ones, the rest is 0.
  rvk::Cuda2DTransferFunction tf;
  auto r2 = rect(4, 9, 0, 5);
  r2->SetColour(1, 1, 1);
  r2->SetOpacity(1);
  const uint64_t m0 = tf.GetMTime();
  tf.AddFunctionObject(r);
  tf.AddFunctionObject(r2);
  ok = tf.GetMTime() > m0 && tf.GetNumberOfFunctionObjects() == 2;
  ok &= tf.getMinIntensity() == 2 && tf.getMaxIntensity() == 9 && tf.getMinGradient() == 0 && tf.getMaxGradient() == 8;
  std::vector<float> a2(n * n, -1);
  tf.GetTransferTable(nullptr, nullptr, nullptr, a2.data(), n, n, 0, 10, 0, 0, 10, 0, 0);
  ok &= a2[3 + 4 * n] == 0.125f && a2[5 + 4 * n] == 1.0f && a2[0] == 0.0f && a2[7 + 2 * n] == 1.0f;
  tf.RemoveFunctionObject(r2);
  ok &= tf.GetNumberOfFunctionObjects() == 1;
  report("composite_2d_function", ok);
}

void cpu_2tf_roundtrip() {
  std::vector<std::shared_ptr<rvk::FunctionPolygon>> in, out;
  auto a = rect(1.5f, 100.25f, 0.0f, 33.3f);
  a->SetColour(0.1f, 0.2f, 0.3f);
  a->SetOpacity(0.4f);
  a->SetAmbient(0.5f);
  a->SetDiffuse(0.6f);
  a->SetSpecular(0.7f);
  a->SetSpecularPower(12.5f);
  a->SetIdentifier(7);
  auto b = std::make_shared<rvk::FunctionPolygon>();
  b->AddVertex(-3.25f, 1e-3f);
  b->AddVertex(1e4f, 2.0f / 3.0f);
  b->AddVertex(5, 6);
  b->SetIdentifier(-2);
  in = {a, b};
  const char* tmp = std::getenv("TMPDIR");
  const std::string path = std::string(tmp ? tmp : "/tmp") + "/rvk_mapper_test_" + std::to_string(getpid()) + ".2tf";
  bool ok = rvk::WriteFunctionPolygons(path, in) && rvk::ReadFunctionPolygons(path, out);
  ok &= out.size() == 2;
  for (size_t k = 0; ok && k < 2; ++k) {
    const auto& p = *in[k];
    const auto& q = *out[k];
    ok &= p.GetRedColourValue() == q.GetRedColourValue() && p.GetGreenColourValue() == q.GetGreenColourValue() &&
          p.GetBlueColourValue() == q.GetBlueColourValue() && p.GetOpacity() == q.GetOpacity() &&
          p.GetAmbient() == q.GetAmbient() && p.GetDiffuse() == q.GetDiffuse() &&
          p.GetSpecular() == q.GetSpecular() && p.GetSpecularPower() == q.GetSpecularPower() &&
          p.GetIdentifier() == q.GetIdentifier() && p.GetNumberOfVertices() == q.GetNumberOfVertices();
    for (size_t v = 0; ok && v < p.GetNumberOfVertices(); ++v)
      ok &= p.GetVertexIntensity(v) == q.GetVertexIntensity(v) && p.GetVertexGradient(v) == q.GetVertexGradient(v);
  }
  std::remove(path.c_str());
  // the text layout itself: count, then colour/opacity, shading, id, vertices
  const std::string p2 = path + ".b";
  ok &= rvk::WriteFunctionPolygons(p2, {rect(0, 1, 2, 3)});
  FILE* f = std::fopen(p2.c_str(), "r");
  char buf[512] = {0};
  size_t got = f ? std::fread(buf, 1, sizeof(buf) - 1, f) : 0;
  if (f) std::fclose(f);
  std::remove(p2.c_str());
  ok &= got > 0 && std::string(buf) == "1\n0 0 0 0\n1 0 0 1\n0\n4\n0 2\n1 2\n1 3\n0 3\n";
  report("2tf_roundtrip", ok, buf);
  ok = !rvk::ReadFunctionPolygons(path + ".missing", out) && out.empty();
  report("2tf_missing_file", ok);
}

void cpu_mapper_state() {
  rvk::CudaDRRImageVolumeMapper drr;
  bool ok = drr.GetCTIntercept() == 0.022f && drr.GetCTSlope() == 0.000022f && drr.GetCTOffset() == 1.0f;
  const double tint[4] = {0.5, 1.5, -1, 1};
  drr.SetTint(tint);
  unsigned char t8[4];
  drr.GetTint(t8);
  ok &= t8[0] == 127 && t8[1] == 255 && t8[2] == 0 && t8[3] == 255;
  report("drr_defaults_and_tint", ok);

  rvk::Cuda1DVolumeMapper m;
  m.SetNumberOfFrames(0);
  ok = m.HasError() && m.GetNumberOfFrames() == 1;
  m.SetNumberOfFrames(31);
  ok &= m.GetNumberOfFrames() == 1;
  rvk::Cuda1DVolumeMapper m2;
  rvk::Renderer ren;
  rvk::Volume vol;
  m2.Render(ren, vol);  // no input
  ok &= m2.HasError() && m2.GetLastError().find("without input") != std::string::npos;
  m2.ChangeFrame(5);  // out of range: ignored
  ok &= m2.GetCurrentFrame() == 0;
  report("mapper_error_latch", ok, m2.GetLastError());

  // component counts: the dual-image mapper takes 2, the others 1; both are
  // refused before any device work
  std::vector<uint8_t> vox(4 * 4 * 4 * 2, 7);
  rvk::ImageData im;
  im.dims[0] = im.dims[1] = im.dims[2] = 4;
  im.scalars = vox.data();
  rvk::CudaDualImageVolumeMapper dual;
  dual.SetInputData(im);
  ok = dual.HasError() && dual.GetLastError().find("2 components") != std::string::npos && !dual.GetInput();
  im.components = 2;
  rvk::Cuda2DInExLogicVolumeMapper inex;
  inex.SetInputData(im);
  ok &= inex.HasError() && inex.GetLastError().find("1 scalar component") != std::string::npos;
  double r1[2], r2[2];
  vox[1] = 9;
  vox[2] = 3;
  im.GetScalarRange(r1, 0);
  im.GetScalarRange(r2, 1);
  ok &= r1[0] == 3 && r1[1] == 7 && r2[0] == 7 && r2[1] == 9;
  report("mapper_component_counts", ok, dual.GetLastError() + " / " + inex.GetLastError());
}

// ---------------------------------------------------------------------------------
struct Scene {
  std::vector<uint8_t> vox[2];
  rvk::ImageData im[2];
  rvk::Renderer ren;
};

void make_scene(Scene& s, int W, int H) {
  const int32_t d[3] = {48, 40, 36};
  for (int f = 0; f < 2; ++f) {
    s.vox[f].resize((size_t)d[0] * d[1] * d[2]);
    rvk_synth_blobs(d, 11 + f, 6, 80, 220, 3, 8, 10, 255, 0, s.vox[f].data());
    rvk::ImageData& im = s.im[f];
    im.dims[0] = d[0];
    im.dims[1] = d[1];
    im.dims[2] = d[2];
    im.spacing[0] = 1.0;
    im.spacing[1] = 1.25;
    im.spacing[2] = 0.9;
    im.origin[0] = -10;
    im.origin[1] = 5;
    im.origin[2] = 2;
    im.scalar_type = RVK_U8;
    im.scalars = s.vox[f].data();
  }
  s.ren.size[0] = W;
  s.ren.size[1] = H;
  rvk::Camera& c = s.ren.camera;
  const double ctr[3] = {-10 + 23.5, 5 + 39 * 1.25 / 2, 2 + 35 * 0.9 / 2};
  for (int a = 0; a < 3; ++a) c.focal_point[a] = ctr[a];
  c.position[0] = ctr[0] + 40;
  c.position[1] = ctr[1] - 30;
  c.position[2] = ctr[2] + 110;
  c.view_up[0] = 0;
  c.view_up[1] = 1;
  c.view_up[2] = 0;
  c.view_angle = 30;
  c.clipping_range[0] = 1;
  c.clipping_range[1] = 500;
}

template <typename T> bool same(const std::vector<T>& a, const std::vector<T>& b, std::string& why) {
  if (a.size() != b.size()) {
    why = fmt("size %g vs %g", (double)a.size(), (double)b.size());
    return false;
  }
  size_t bad = 0;
  for (size_t i = 0; i < a.size(); ++i) bad += !(a[i] == b[i]);
  if (bad) why = fmt("%g of %g elements differ", (double)bad, (double)a.size());
  return bad == 0;
}

bool lsb_close(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b, std::string& why) {
  if (a.size() != b.size() || a.empty()) {
    why = "size";
    return false;
  }
  size_t over1 = 0, ne = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    const int dd = std::abs((int)a[i] - (int)b[i]);
    over1 += dd > 1;
    ne += dd != 0;
  }
  why = fmt("%g channels > 1 LSB, %g differ of %g", (double)over1, (double)ne, (double)a.size());
  return over1 == 0 && ne * 1000 <= a.size();
}

struct DeviceImage {
  rvk_ctx* ctx = nullptr;
  void* p = nullptr;
  size_t n = 0;
  DeviceImage(rvk_ctx* c, size_t bytes) : ctx(c), n(bytes) { rvk_device_alloc(ctx, bytes, &p); }
  ~DeviceImage() {
    if (p) rvk_device_free(ctx, p);
  }
  template <typename T> std::vector<T> get() {
    std::vector<T> h(n / sizeof(T));
    rvk_copy_to_host(ctx, h.data(), p, n);
    return h;
  }
};

void gpu_drr(Scene& s) {
  rvk::CudaDRRImageVolumeMapper m;
  m.SetInputData(s.im[0]);
  m.SetOutputMode(RVK_DRR_F32_EXP_NEG_L);
  m.SetCTIntercept(0.0f);
  m.SetCTSlope(0.002f);
  m.SetSkipRule(false);
  rvk::Volume vol;
  m.Render(s.ren, vol);
  if (m.HasError()) {
    report("drr_mapper_render", false, m.GetLastError());
    return;
  }
  // the mapper's view equals rvk_compute_view of the same camera
  rvk_view v;
  const rvk_camera cam = s.ren.camera.ToABI();
  rvk_compute_view(&cam, nullptr, s.im[0].origin, s.im[0].spacing, s.ren.size[0], s.ren.size[1], 0, &v);
  bool ok = std::memcmp(v.view_to_voxels, m.GetView().view_to_voxels, sizeof(v.view_to_voxels)) == 0 &&
            m.GetView().width == s.ren.size[0] && m.GetView().height == s.ren.size[1];
  report("drr_mapper_view", ok);

  rvk_ctx* ctx = m.GetContext();
  rvk_volume* vo = nullptr;
  const int32_t d[3] = {s.im[0].dims[0], s.im[0].dims[1], s.im[0].dims[2]};
  rvk_volume_create(ctx, d, s.im[0].spacing, RVK_U8, 1, &vo);
  rvk_volume_set_frame(vo, 0, s.vox[0].data(), 0);
  rvk_march_params mp{1.0f, 0.0f, 1, 0, 0};
  rvk_drr_params dp{0.0f, 0.002f, 1.0f, RVK_DRR_F32_EXP_NEG_L, 1, {0, 0, 0, 0}};
  const size_t px = (size_t)v.width * v.height;
  DeviceImage out(ctx, px * 4);
  int rc = rvk_render_drr(ctx, vo, 0, &m.GetView(), &mp, &dp, out.p);
  std::string why;
  ok = rc == RVK_OK && same(out.get<float>(), m.GetOutput().values, why);
  float lo = 1, hi = 0;
  for (float x : m.GetOutput().values) {
    lo = std::min(lo, x);
    hi = std::max(hi, x);
  }
  ok &= lo < 0.9f && hi == 1.0f;  // some rays attenuate, rays that miss keep exp(0)
  report("drr_mapper_vs_direct", ok, why + fmt(" range %g..%g", lo, hi));

  // the reference's RGBA8 output with a tint, and a half-resolution render
  m.SetOutputMode(RVK_DRR_RGBA8_REF);
  const unsigned char tint[4] = {255, 128, 0, 200};
  m.SetTint(tint);
  m.SetRenderOutputScaleFactor(2.0f);
  m.Render(s.ren, vol);
  rvk_view vh = m.GetView();
  ok = !m.HasError() && vh.width == s.ren.size[0] / 2 && vh.height == s.ren.size[1] / 2;
  dp.output = RVK_DRR_RGBA8_REF;
  std::memcpy(dp.tint, tint, 4);
  DeviceImage out8(ctx, (size_t)vh.width * vh.height * 4);
  rc = rvk_render_drr(ctx, vo, 0, &vh, &mp, &dp, out8.p);
  ok &= rc == RVK_OK && same(out8.get<uint8_t>(), m.GetOutput().rgba, why);
  report("drr_mapper_rgba_half_resolution", ok, why);
  rvk_volume_destroy(vo);
}

void gpu_1d(Scene& s) {
  rvk::Cuda1DVolumeMapper m;
  m.SetNumberOfFrames(2);
  m.SetInputData(s.im[0], 0);
  m.SetInputData(s.im[1], 1);
  m.ChangeFrame(1);
  rvk::Volume vol;
  double sr[2];
  s.im[1].GetScalarRange(sr);
  vol.property.scalar_opacity.AddPoint(sr[0], 0.0);
  vol.property.scalar_opacity.AddPoint(sr[1], 1.0);
  vol.property.color.AddRGBPoint(sr[0], 1, 0, 0);
  vol.property.color.AddRGBPoint(sr[1], 0, 0, 1);
  vol.property.shade = true;
  vol.property.ambient = 0.3;
  vol.property.diffuse = 0.6;
  vol.property.specular = 0.2;
  vol.property.specular_power = 8;
  const int n = 256;
  m.SetTransferFunctionSize(n);
  m.Render(s.ren, vol);
  if (m.HasError()) {
    report("1d_mapper_render", false, m.GetLastError());
    return;
  }
  // direct render with the tables written in closed form (linear ramps over
  // the scalar range) and the same post shading
  rvk_ctx* ctx = m.GetContext();
  rvk_volume* vo = nullptr;
  const int32_t d[3] = {s.im[1].dims[0], s.im[1].dims[1], s.im[1].dims[2]};
  rvk_volume_create(ctx, d, s.im[1].spacing, RVK_U8, 1, &vo);
  rvk_volume_set_frame(vo, 0, s.vox[1].data(), 0);
  std::vector<float> a(n), r(n), g(n, 0.0f), b(n);
  for (int i = 0; i < n; ++i) {
    a[i] = (float)((double)i / (n - 1));
    r[i] = 1.0f - a[i];
    b[i] = a[i];
  }
  rvk_tf1d_params p{(float)sr[0], (float)(1.0 / (sr[1] - sr[0])), 0.0f, INFINITY, 0.3f, 0.6f, 0.2f, 8.0f};
  rvk_tf1d* tf = nullptr;
  rvk_tf1d_create(ctx, n, a.data(), nullptr, r.data(), g.data(), b.data(), &p, &tf);
  const rvk_view& v = m.GetView();
  const size_t px = (size_t)v.width * v.height;
  DeviceImage out(ctx, px * 4), dep(ctx, px * 4), mn(ctx, px * 4), mx(ctx, px * 4);
  rvk_march_params mp{1.0f, 1.0f / 64.0f, 1, 1, 0};
  rvk_aux_out aux{(float*)dep.p, (float*)mn.p, (float*)mx.p};
  int rc = rvk_render_1d(ctx, vo, 0, tf, &v, &mp, (uint8_t*)out.p, &aux);
  rvk_shade_params sh{0.70f, 0.002f, (float)(1.0 / (0.005 - 0.002)), 0.25f, 0.45f, (float)(1.0 / (0.6 - 0.45))};
  if (rc == RVK_OK) rc = rvk_post_shade(ctx, v.width, v.height, &sh, aux.depth, aux.min_depth, aux.max_depth, (uint8_t*)out.p);
  std::string why;
  const auto direct = out.get<uint8_t>();
  bool ok = rc == RVK_OK && lsb_close(direct, m.GetOutput().rgba, why);
  size_t lit = 0;
  for (size_t i = 3; i < direct.size(); i += 4) lit += direct[i] > 0;
  ok &= lit > px / 20;
  report("1d_mapper_frame1_vs_direct", ok, why + fmt(" lit %g", (double)lit));
  rvk_tf1d_destroy(tf);

  // a device move with data: the cached frames are uploaded again
  const std::vector<uint8_t> before = m.GetOutput().rgba;
  m.SetDevice(0, true);
  m.Render(s.ren, vol);
  ok = !m.HasError() && same(before, m.GetOutput().rgba, why);
  report("1d_mapper_set_device_with_data", ok, why + m.GetLastError());

  // an all-far z-buffer changes nothing
  std::vector<float> zb((size_t)s.ren.size[0] * s.ren.size[1], 1.0f);
  rvk::Renderer rz = s.ren;
  rz.zbuffer = zb.data();
  m.Render(rz, vol);
  ok = !m.HasError() && same(before, m.GetOutput().rgba, why);
  report("1d_mapper_far_zbuffer", ok, why + m.GetLastError());
  rvk_volume_destroy(vo);
}

// Six clipping planes around a world-space box: the voxel-space planes the
// mapper derives satisfy n_vox . v + d = n_world . (p - o) for any world point
// p with voxel coordinates v.
void gpu_planes(Scene& s) {
  rvk::CudaDRRImageVolumeMapper m;
  m.SetInputData(s.im[0]);
  m.SetOutputMode(RVK_DRR_F32_L);
  m.SetCTIntercept(0.0f);
  m.SetCTSlope(0.002f);
  rvk::Volume vol;
  vol.has_user_matrix = true;
  const double um[16] = {0.8, -0.6, 0, 3, 0.6, 0.8, 0, -2, 0, 0, 1, 1.5, 0, 0, 0, 1};
  std::memcpy(vol.user_matrix, um, sizeof(um));
  const double ctr[3] = {s.ren.camera.focal_point[0], s.ren.camera.focal_point[1], s.ren.camera.focal_point[2]};
  const double n6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int i = 0; i < 6; ++i) {
    rvk::Plane pl;
    for (int a = 0; a < 3; ++a) {
      pl.normal[a] = n6[i][a];
      pl.origin[a] = ctr[a] - 8.0 * n6[i][a];
    }
    m.AddClippingPlane(pl);
  }
  m.Render(s.ren, vol);
  const rvk_view& v = m.GetView();
  bool ok = !m.HasError() && v.n_clip_planes == 6 && v.n_keyhole_planes == 0;
  // voxel -> world: U * diag(spacing) + origin
  const rvk::ImageData& im = s.im[0];
  double worst = 0;
  const double vx[3][3] = {{3, 7, 11}, {40, 2, 30}, {20, 20, 20}};
  for (const auto& q : vx) {
    double w0[3];
    for (int a = 0; a < 3; ++a) w0[a] = q[a] * im.spacing[a] + im.origin[a];
    double p[3];
    for (int r = 0; r < 3; ++r) p[r] = um[4 * r] * w0[0] + um[4 * r + 1] * w0[1] + um[4 * r + 2] * w0[2] + um[4 * r + 3];
    for (int i = 0; i < 6; ++i) {
      const float* pl = v.clip_planes + 4 * i;
      const double lhs = pl[0] * q[0] + pl[1] * q[1] + pl[2] * q[2] + pl[3];
      double rhs = 0;
      for (int a = 0; a < 3; ++a) rhs += n6[i][a] * (p[a] - (ctr[a] - 8.0 * n6[i][a]));
      worst = std::max(worst, std::fabs(lhs - rhs));
    }
  }
  ok &= worst < 1e-4;
  // clipped image: fewer attenuated pixels than the unclipped one.  A few
  // rays that graze the volume box can attenuate only when clipped: the
  // volume crop keeps the direction of the uncropped ray for all three axes
  // (V/CUDA_vtkCudaVolumeMapper_renderAlgo.cu:199-318), so an unclipped
  // grazing ray can end with start and end swapped and a sub-step length.
  const std::vector<float> clipped = m.GetOutput().values;
  m.RemoveAllClippingPlanes();
  m.Render(s.ren, vol);
  const std::vector<float>& full = m.GetOutput().values;
  size_t nz_clip = 0, nz_full = 0, extra = 0;
  for (size_t i = 0; i < full.size(); ++i) {
    nz_clip += clipped[i] > 0;
    nz_full += full[i] > 0;
    extra += clipped[i] > 0 && full[i] == 0;
  }
  ok &= !m.HasError() && nz_clip > 0 && nz_clip < nz_full && extra * 1000 <= full.size() &&
        clipped.size() == full.size();
  report("clipping_planes_voxel_space", ok,
         fmt("plane residual %g, nonzero %g of %g", worst, (double)nz_clip, (double)nz_full) +
             fmt(", %g grazing", (double)extra));

  // a 5-plane set is not applied (FigurePlanes needs exactly six)
  for (int i = 0; i < 5; ++i) {
    rvk::Plane pl;
    for (int a = 0; a < 3; ++a) pl.normal[a] = n6[i][a];
    m.AddKeyholePlane(pl);
  }
  m.Render(s.ren, vol);
  report("five_planes_ignored", !m.HasError() && m.GetView().n_keyhole_planes == 0);
}

void gpu_2d(Scene& s) {
  rvk::Cuda2DVolumeMapper m;
  m.SetInputData(s.im[0]);
  auto fn = std::make_shared<rvk::Cuda2DTransferFunction>();
  auto poly = rect(60, 250, 0, 400);
  poly->SetColour(0.9f, 0.6f, 0.2f);
  poly->SetOpacity(0.3f);
  poly->SetAmbient(0.4f);
  poly->SetDiffuse(0.5f);
  poly->SetSpecular(0.3f);
  poly->SetSpecularPower(6.0f);
  fn->AddFunctionObject(poly);
  m.SetFunction(fn);
  const int n = 128;
  m.SetTransferFunctionSize(n);
  m.SetPostShading(false);
  rvk::Volume vol;
  m.Render(s.ren, vol);
  if (m.HasError()) {
    report("2d_mapper_render", false, m.GetLastError());
    return;
  }
  // direct render: tables and parameters from the reference's recipe, written
  // out here for a single rectangle
  const rvk::ImageData& im = s.im[0];
  double sr[2];
  im.GetScalarRange(sr);
  const double minI = std::max(sr[0], 60.0), maxI = std::min(sr[1], 250.0);
  const double highG = std::fabs(sr[0] - sr[1]) / 0.9;
  double maxG = std::min(highG, 400.0), minG = 0.0;
  const double goff = maxG * 0.8;
  maxG = (std::log(maxG * maxG + goff) - std::log(goff)) / std::log(2.0) + 1.0;
  minG = (std::log(minG * minG + goff) - std::log(goff)) / std::log(2.0);
  rvk_tf2d_params p{(float)minI, (float)(1.0 / (maxI - minI)), (float)(-minG - std::log(goff) / std::log(2.0)),
                    (float)(1.0 / (maxG - minG)), (float)goff};
  std::vector<std::vector<float>> t(8, std::vector<float>(n * n, 0.0f));
  const float vals[8] = {0.3f, 0.9f, 0.6f, 0.2f, 0.4f, 0.5f, 0.3f, 6.0f};
  // rows: bounding rows [lg, hg) whose sampled gradient lies in (0, 400];
  // columns: [li, hi) whose sampled intensity lies in (60, 250]
  const float fminI = (float)minI, fmaxI = (float)maxI, fminG = (float)minG, fmaxG = (float)maxG, fgoff = (float)goff;
  auto idx = [&](float v, float lo, float hi) {
    int k = (int)((double)(n - 1) * (v - lo) / (hi - lo));
    return std::min(std::max(k, 0), n - 1);
  };
  auto lv = [&](float gr) { return (float)((std::log((double)(gr * gr + fgoff)) - std::log((double)fgoff)) / std::log(2.0)); };
  const int li = idx(60.0f, fminI, fmaxI), hi = idx(250.0f, fminI, fmaxI);
  const int lg = idx(lv(0.0f), fminG, fmaxG), hg = idx(lv(400.0f), fminG, fmaxG);
  for (int g = lg; g < hg; ++g) {
    const double v = (double)g * (double)(fmaxG - fminG) / (double)(n - 1) + fminG;
    const float gr = (float)std::sqrt(std::exp(std::log(2.0) * v + std::log((double)fgoff)) - fgoff);
    if (!(gr > 0.0f && gr <= 400.0f)) continue;
    for (int i = li; i < hi; ++i) {
      const float in = (float)((double)i * (fmaxI - fminI) / (double)(n - 1) + fminI);
      if (!(in > 60.0f && in <= 250.0f)) continue;
      for (int k = 0; k < 8; ++k) t[k][i + g * n] = vals[k];
    }
  }
  rvk_ctx* ctx = m.GetContext();
  rvk_volume* vo = nullptr;
  const int32_t d[3] = {im.dims[0], im.dims[1], im.dims[2]};
  rvk_volume_create(ctx, d, im.spacing, RVK_U8, 1, &vo);
  rvk_volume_set_frame(vo, 0, s.vox[0].data(), 0);
  const float* tp[8];
  for (int k = 0; k < 8; ++k) tp[k] = t[k].data();
  rvk_tf2d* tf = nullptr;
  rvk_tf2d_create(ctx, n, tp, nullptr, &p, &tf);
  const rvk_view& v = m.GetView();
  const size_t px = (size_t)v.width * v.height;
  DeviceImage out(ctx, px * 4);
  rvk_march_params mp{1.0f, 1.0f / 32.0f, 1, 1, 0};
  int rc = rvk_render_2d(ctx, vo, 0, tf, &v, &mp, 0, (uint8_t*)out.p, nullptr);
  std::string why;
  const auto direct = out.get<uint8_t>();
  bool ok = rc == RVK_OK && lsb_close(direct, m.GetOutput().rgba, why);
  size_t lit = 0;
  for (size_t i = 3; i < direct.size(); i += 4) lit += direct[i] > 0;
  ok &= lit > px / 20;
  report("2d_mapper_polygon_vs_direct", ok, why + fmt(" lit %g", (double)lit));

  // precomputed gradient-magnitude volume path renders without error and
  // stays close to the six-gather path in coverage
  m.SetGradientMagnitudeVolume(400.0f);
  m.Render(s.ren, vol);
  size_t lit2 = 0;
  for (size_t i = 3; i < m.GetOutput().rgba.size(); i += 4) lit2 += m.GetOutput().rgba[i] > 0;
  ok = !m.HasError() && lit2 > lit / 2;
  report("2d_mapper_gradmag_volume", ok, m.GetLastError() + fmt(" lit %g vs %g", (double)lit2, (double)lit));
  rvk_tf2d_destroy(tf);
  rvk_volume_destroy(vo);
}

// The 2D recipe's scalars (intensity range from the function and the input,
// log-scaled gradient range clipped at |range| / min spacing) for a function
// spanning intensities [i0, i1] and gradients [0, g1].
rvk_tf2d_params recipe_2d(const rvk::ImageData& im, float i0, float i1, float g1, float out[6]) {
  double sr[2];
  im.GetScalarRange(sr);
  const double minI = std::max(sr[0], (double)i0), maxI = std::min(sr[1], (double)i1);
  const double highG = std::fabs(sr[0] - sr[1]) / 0.9;
  double maxG = std::min(highG, (double)g1), minG = 0.0;
  const double goff = maxG * 0.8;
  maxG = (std::log(maxG * maxG + goff) - std::log(goff)) / std::log(2.0) + 1.0;
  minG = (std::log(minG * minG + goff) - std::log(goff)) / std::log(2.0);
  const float o[6] = {(float)minI, (float)maxI, 0.0f, (float)minG, (float)maxG, (float)goff};
  std::memcpy(out, o, sizeof(o));
  return rvk_tf2d_params{(float)minI, (float)(1.0 / (maxI - minI)), (float)(-minG - std::log(goff) / std::log(2.0)),
                         (float)(1.0 / (maxG - minG)), (float)goff};
}

void gpu_inex(Scene& s) {
  rvk::Cuda2DInExLogicVolumeMapper m;
  m.SetInputData(s.im[0]);
  auto fn = std::make_shared<rvk::Cuda2DTransferFunction>();
  auto poly = rect(110, 250, 0, 400);
  poly->SetColour(0.3f, 0.8f, 0.5f);
  poly->SetOpacity(0.25f);
  poly->SetAmbient(0.5f);
  poly->SetDiffuse(0.5f);
  poly->SetSpecularPower(1.0f);
  fn->AddFunctionObject(poly);
  // a transparent polygon widens the table's intensity range down to 80
  fn->AddFunctionObject(rect(80, 110, 0, 400));
  // in/ex: exclude rays that meet the bright core first, mark the shell
  auto ie = std::make_shared<rvk::Cuda2DTransferFunction>();
  auto core = rect(170, 250, 0, 400);
  core->SetOpacity(1.0f);
  auto shell = rect(90, 110, 0, 400);
  shell->SetOpacity(0.3f);
  ie->AddFunctionObject(core);
  ie->AddFunctionObject(shell);
  m.SetVisualizationFunction(fn);
  m.SetInExLogicFunction(ie);
  m.SetUseBlackKeyhole(true);
  const int n = 128;
  m.SetTransferFunctionSize(n);
  m.SetPostShading(false);
  rvk::Volume vol;
  m.Render(s.ren, vol);
  if (m.HasError()) {
    report("inex_mapper_render", false, m.GetLastError());
    return;
  }
  const rvk::ImageData& im = s.im[0];
  float a[6];
  const rvk_tf2d_params p = recipe_2d(im, 80, 250, 400, a);
  std::vector<std::vector<float>> t(8, std::vector<float>(n * n));
  std::vector<float> inex(n * n);
  fn->GetTransferTable(t[1].data(), t[2].data(), t[3].data(), t[0].data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5], 2);
  fn->GetShadingTable(t[4].data(), t[5].data(), t[6].data(), t[7].data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5], 2);
  ie->GetTransferTable(nullptr, nullptr, nullptr, inex.data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5], 2);
  rvk_ctx* ctx = m.GetContext();
  rvk_volume* vo = nullptr;
  const int32_t d[3] = {im.dims[0], im.dims[1], im.dims[2]};
  rvk_volume_create(ctx, d, im.spacing, RVK_U8, 1, &vo);
  rvk_volume_set_frame(vo, 0, s.vox[0].data(), 0);
  const float* tp[8];
  for (int k = 0; k < 8; ++k) tp[k] = t[k].data();
  rvk_tf2d* tf = nullptr;
  rvk_tf2d_create(ctx, n, tp, nullptr, &p, &tf);
  rvk_tf2d_set_inex(tf, inex.data());
  const rvk_view& v = m.GetView();
  const size_t px = (size_t)v.width * v.height;
  DeviceImage out(ctx, px * 4);
  rvk_march_params mp{1.0f, 1.0f / 32.0f, 1, 1, 0};
  int rc = rvk_render_2d_inex(ctx, vo, 0, tf, &v, &mp, 1, (uint8_t*)out.p, nullptr);
  std::string why;
  const auto direct = out.get<uint8_t>();
  bool ok = rc == RVK_OK && lsb_close(direct, m.GetOutput().rgba, why);
  size_t black = 0, lit = 0;
  for (size_t i = 0; i < direct.size(); i += 4) {
    lit += direct[i + 3] > 0;
    black += direct[i + 3] == 255 && direct[i] == 0 && direct[i + 1] == 0 && direct[i + 2] == 0;
  }
  ok &= lit > px / 50 && black > 0 && lit < px;
  report("inex_mapper_polygon_vs_direct", ok, why + fmt(" lit %g black %g", (double)lit, (double)black));
  rvk_tf2d_destroy(tf);
  rvk_volume_destroy(vo);
}

void gpu_dual(Scene& s) {
  const rvk::ImageData& im1 = s.im[0];
  const size_t nv = s.vox[0].size();
  std::vector<uint8_t> both(2 * nv);
  for (size_t i = 0; i < nv; ++i) {
    both[2 * i] = s.vox[0][i];
    both[2 * i + 1] = s.vox[1][i];
  }
  rvk::ImageData im = im1;
  im.components = 2;
  im.scalars = both.data();
  rvk::CudaDualImageVolumeMapper m;
  m.SetInputData(im);
  auto fn = std::make_shared<rvk::Cuda2DTransferFunction>();
  auto poly = rect(60, 250, 60, 255);
  poly->SetColour(0.9f, 0.4f, 0.7f);
  poly->SetOpacity(0.2f);
  poly->SetAmbient(0.6f);
  poly->SetDiffuse(0.4f);
  poly->SetSpecular(0.2f);
  poly->SetSpecularPower(4.0f);
  fn->AddFunctionObject(poly);
  m.SetFunction(fn);
  const int n = 128;
  m.SetTransferFunctionSize(n);
  m.SetPostShading(false);
  rvk::Volume vol;
  m.Render(s.ren, vol);
  if (m.HasError()) {
    report("dual_mapper_render", false, m.GetLastError());
    return;
  }
  double sr[4];
  im.GetScalarRange(sr, 0);
  im.GetScalarRange(sr + 2, 1);
  const double lo1 = std::max(sr[0], 60.0), hi1 = std::min(sr[1], 250.0);
  const double lo2 = std::max(sr[2], 60.0), hi2 = std::min(sr[3], 255.0);
  const rvk_dual_params dp{(float)lo1, (float)(1.0 / (hi1 - lo1)), (float)lo2, (float)(1.0 / (hi2 - lo2))};
  std::vector<std::vector<float>> t(8, std::vector<float>(n * n));
  fn->GetTransferTable(t[1].data(), t[2].data(), t[3].data(), t[0].data(), n, n, (float)lo1, (float)hi1, 0.0f,
                       (float)lo2, (float)hi2, 0.0f, 0);
  fn->GetShadingTable(t[4].data(), t[5].data(), t[6].data(), t[7].data(), n, n, (float)lo1, (float)hi1, 0.0f,
                      (float)lo2, (float)hi2, 0.0f, 0);
  rvk_ctx* ctx = m.GetContext();
  rvk_volume *v1 = nullptr, *v2 = nullptr;
  const int32_t d[3] = {im.dims[0], im.dims[1], im.dims[2]};
  rvk_volume_create(ctx, d, im.spacing, RVK_U8, 1, &v1);
  rvk_volume_create(ctx, d, im.spacing, RVK_U8, 1, &v2);
  rvk_volume_set_frame(v1, 0, s.vox[0].data(), 0);
  rvk_volume_set_frame(v2, 0, s.vox[1].data(), 0);
  const float* tp[8];
  for (int k = 0; k < 8; ++k) tp[k] = t[k].data();
  const rvk_tf2d_params unused{};
  rvk_tf2d* tf = nullptr;
  rvk_tf2d_create(ctx, n, tp, nullptr, &unused, &tf);
  const rvk_view& v = m.GetView();
  const size_t px = (size_t)v.width * v.height;
  DeviceImage out(ctx, px * 4);
  rvk_march_params mp{1.0f, 1.0f / 32.0f, 1, 1, 0};
  int rc = rvk_render_dual(ctx, v1, v2, 0, tf, &dp, &v, &mp, (uint8_t*)out.p, nullptr);
  std::string why;
  const auto direct = out.get<uint8_t>();
  bool ok = rc == RVK_OK && lsb_close(direct, m.GetOutput().rgba, why);
  size_t lit = 0;
  for (size_t i = 3; i < direct.size(); i += 4) lit += direct[i] > 0;
  ok &= lit > px / 50;
  // swapping the components changes the picture (the axes are not symmetric)
  int rc2 = rvk_render_dual(ctx, v2, v1, 0, tf, &dp, &v, &mp, (uint8_t*)out.p, nullptr);
  std::string why2;
  ok &= rc2 == RVK_OK && !lsb_close(out.get<uint8_t>(), direct, why2);
  report("dual_mapper_polygon_vs_direct", ok, why + fmt(" lit %g", (double)lit));
  rvk_tf2d_destroy(tf);
  rvk_volume_destroy(v1);
  rvk_volume_destroy(v2);
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "cpu";
  if (mode == "cpu") {
    cpu_piecewise();
    cpu_polygons();
    cpu_2tf_roundtrip();
    cpu_mapper_state();
  } else if (mode == "gpu") {
    int nd = 0;
    if (rvk_device_count(&nd) != RVK_OK || nd < 1) {
      std::printf("FAIL no device: %s\n", rvk_last_error_string());
      return 1;
    }
    Scene s;
    make_scene(s, 200, 150);
    gpu_drr(s);
    gpu_1d(s);
    gpu_planes(s);
    gpu_2d(s);
    gpu_inex(s);
    gpu_dual(s);
  } else {
    std::fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
    return 2;
  }
  std::printf("%d failures\n", g_fail);
  return std::min(g_fail, 100);
}
