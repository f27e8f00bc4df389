// rvk_mapper.cpp -- the VTK-free mapper classes declared in rvk_mapper.h.
//
// Every device operation goes through the librvkhip C-ABI.  Paths below are
// relative to the reference root; V/ = Visualization/vtkCudaVisualization,
// C/ = Common/vtkCudaCommon.
#include "rvk_mapper.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>

namespace rvk {

namespace {

template <typename T> void range_of(const T* d, size_t n, double r[2], int stride = 1) {
  if (!n) { r[0] = r[1] = 0.0; return; }
  T lo = d[0], hi = d[0];
  for (size_t i = 1; i < n; ++i) {
    lo = std::min(lo, d[i * stride]);
    hi = std::max(hi, d[i * stride]);
  }
  r[0] = (double)lo;
  r[1] = (double)hi;
}

size_t dtype_bytes(int t) {
  switch (t) {
    case RVK_U8: case RVK_I8: return 1;
    case RVK_U16: case RVK_I16: return 2;
    case RVK_U32: case RVK_I32: case RVK_F32: return 4;
    default: return 0;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
void ImageData::GetScalarRange(double r[2], int c) const {
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  r[0] = r[1] = 0.0;
  if (!scalars || c < 0 || c >= components) return;
  const int k = components;
  switch (scalar_type) {
    case RVK_U8: range_of((const uint8_t*)scalars + c, n, r, k); break;
    case RVK_I8: range_of((const int8_t*)scalars + c, n, r, k); break;
    case RVK_U16: range_of((const uint16_t*)scalars + c, n, r, k); break;
    case RVK_I16: range_of((const int16_t*)scalars + c, n, r, k); break;
    case RVK_U32: range_of((const uint32_t*)scalars + c, n, r, k); break;
    case RVK_I32: range_of((const int32_t*)scalars + c, n, r, k); break;
    case RVK_F32: range_of((const float*)scalars + c, n, r, k); break;
    default: break;
  }
}

rvk_camera Camera::ToABI() const {
  rvk_camera c;
  std::memset(&c, 0, sizeof(c));
  for (int a = 0; a < 3; ++a) {
    c.position[a] = position[a];
    c.focal_point[a] = focal_point[a];
    c.view_up[a] = view_up[a];
  }
  c.view_angle = view_angle;
  c.clipping_range[0] = clipping_range[0];
  c.clipping_range[1] = clipping_range[1];
  c.parallel_projection = parallel_projection;
  c.parallel_scale = parallel_scale;
  return c;
}

// ---- piecewise functions (vtkPiecewiseFunction / vtkColorTransferFunction) ---
void PiecewiseFunction::AddPoint(double x, double y) {
  auto it = std::lower_bound(pts_.begin(), pts_.end(), x,
                             [](const std::array<double, 2>& p, double v) { return p[0] < v; });
  if (it != pts_.end() && (*it)[0] == x) (*it)[1] = y;
  else pts_.insert(it, {x, y});
}

void PiecewiseFunction::RemoveAllPoints() { pts_.clear(); }

void PiecewiseFunction::GetRange(double& lo, double& hi) const {
  lo = pts_.empty() ? 0.0 : pts_.front()[0];
  hi = pts_.empty() ? 0.0 : pts_.back()[0];
}

double PiecewiseFunction::GetValue(double x) const {
  if (pts_.empty()) return 0.0;
  if (x <= pts_.front()[0]) return pts_.front()[1];
  if (x >= pts_.back()[0]) return pts_.back()[1];
  auto it = std::upper_bound(pts_.begin(), pts_.end(), x,
                             [](double v, const std::array<double, 2>& p) { return v < p[0]; });
  const auto& b = *it;
  const auto& a = *(it - 1);
  const double t = (x - a[0]) / (b[0] - a[0]);
  return a[1] + t * (b[1] - a[1]);
}

void PiecewiseFunction::GetTable(double lo, double hi, int n, float* out) const {
  for (int i = 0; i < n; ++i) {
    const double x = (n > 1) ? lo + (double)i * (hi - lo) / (double)(n - 1) : lo;
    out[i] = (float)GetValue(x);
  }
}

void ColorTransferFunction::AddRGBPoint(double x, double r, double g, double b) {
  r_.AddPoint(x, r);
  g_.AddPoint(x, g);
  b_.AddPoint(x, b);
}

void ColorTransferFunction::RemoveAllPoints() {
  r_.RemoveAllPoints();
  g_.RemoveAllPoints();
  b_.RemoveAllPoints();
}

void ColorTransferFunction::GetRange(double& lo, double& hi) const { r_.GetRange(lo, hi); }

void ColorTransferFunction::GetTable(double lo, double hi, int n, float* rgb) const {
  std::vector<float> t(n);
  const PiecewiseFunction* ch[3] = {&r_, &g_, &b_};
  for (int c = 0; c < 3; ++c) {
    ch[c]->GetTable(lo, hi, n, t.data());
    for (int i = 0; i < n; ++i) rgb[3 * i + c] = t[i];
  }
}

// ---- 2D transfer functions -----------------------------------------------------
void FunctionPolygon::AddVertex(float i, float g) { contour_.push_back({i, g}); }
void FunctionPolygon::ModifyVertex(size_t k, float i, float g) {
  if (k < contour_.size()) contour_[k] = {i, g};
}
void FunctionPolygon::RemoveVertex(size_t k) {
  if (k < contour_.size()) contour_.erase(contour_.begin() + k);
}

double FunctionPolygon::getMinIntensity() const {
  double m = std::numeric_limits<double>::max();
  for (const auto& v : contour_) m = std::min(m, (double)v[0]);
  return m;
}
double FunctionPolygon::getMaxIntensity() const {
  double m = -std::numeric_limits<double>::max();
  for (const auto& v : contour_) m = std::max(m, (double)v[0]);
  return m;
}
double FunctionPolygon::getMinGradient() const {
  double m = std::numeric_limits<double>::max();
  for (const auto& v : contour_) m = std::min(m, (double)v[1]);
  return m;
}
double FunctionPolygon::getMaxGradient() const {
  double m = -std::numeric_limits<double>::max();
  for (const auto& v : contour_) m = std::max(m, (double)v[1]);
  return m;
}

// Even-odd rule: a horizontal ray (constant gradient y) toward -intensity
// crosses an edge when the edge spans y (half-open in gradient) and the
// crossing lies left of x (C/vtkCudaFunctionPolygon.cxx:179-198).
bool FunctionPolygon::PointInPolygon(float x, float y) const {
  const size_t n = contour_.size();
  if (n < 3) return false;
  bool odd = false;
  size_t j = n - 1;
  for (size_t i = 0; i < n; ++i) {
    const float gi = contour_[i][1], gj = contour_[j][1];
    if ((gi < y && gj >= y) || (gj < y && gi >= y)) {
      const float ii = contour_[i][0], ij = contour_[j][0];
      if (ii + (y - gi) / (gj - gi) * (ij - ii) < x) odd = !odd;
    }
    j = i;
  }
  return odd;
}

// Table rasterisation (C/vtkCudaFunctionPolygon.cxx:200-435): the polygon's
// bounding box in table indices [low, high) -- the last row/column of the box
// is never written -- sampled at lo + k (hi - lo)/(size - 1) on each axis,
// the log axes mapped back through sqrt(2^v * off - off).
template <typename F>
void FunctionPolygon::Rasterise(int sizeI, int sizeG, float lowI, float highI, float offI, float lowG, float highG,
                                float offG, int logUsed, F&& write) const {
  float minI = (float)getMinIntensity(), maxI = (float)getMaxIntensity();
  float minG = (float)getMinGradient(), maxG = (float)getMaxGradient();
  // the squares are summed in float, the logarithms taken in double
  auto to_log = [](float v, float off) {
    return (float)((std::log((double)(v * v + off)) - std::log((double)off)) / std::log(2.0));
  };
  if (logUsed & 2) {
    minG = to_log(minG, offG);
    maxG = to_log(maxG, offG);
  }
  if (logUsed & 1) {
    minI = to_log(minI, offI);
    maxI = to_log(maxI, offI);
  }
  auto index = [](int size, float v, float lo, float hi) {
    int k = (int)((double)(size - 1) * (v - lo) / (hi - lo));
    return std::min(std::max(k, 0), size - 1);
  };
  const int li = index(sizeI, minI, lowI, highI), hi_i = index(sizeI, maxI, lowI, highI);
  const int lg = index(sizeG, minG, lowG, highG), hg = index(sizeG, maxG, lowG, highG);
  for (int g = lg; g < hg; ++g) {
    double grad = (double)g * (double)(highG - lowG) / (double)(sizeG - 1) + lowG;
    if (logUsed & 2) grad = std::sqrt(std::exp(std::log(2.0) * grad + std::log((double)offG)) - offG);
    for (int i = li; i < hi_i; ++i) {
      double in = (double)i * (double)(highI - lowI) / (double)(sizeI - 1) + lowI;
      if (logUsed & 1) in = std::sqrt(std::exp(std::log(2.0) * in + std::log((double)offI)) - offI);
      if (PointInPolygon((float)in, (float)grad)) write(i + g * sizeI);
    }
  }
}

void FunctionPolygon::PopulatePortionOfTransferTable(int sizeI, int sizeG, float lowI, float highI, float offI,
                                                     float lowG, float highG, float offG, float* r, float* g,
                                                     float* b, float* a, int logUsed) const {
  Rasterise(sizeI, sizeG, lowI, highI, offI, lowG, highG, offG, logUsed, [&](int k) {
    if (r) r[k] = r_;
    if (g) g[k] = g_;
    if (b) b[k] = b_;
    if (a) a[k] = a_;
  });
}

void FunctionPolygon::PopulatePortionOfShadingTable(int sizeI, int sizeG, float lowI, float highI, float offI,
                                                    float lowG, float highG, float offG, float* amb, float* dif,
                                                    float* spec, float* pw, int logUsed) const {
  Rasterise(sizeI, sizeG, lowI, highI, offI, lowG, highG, offG, logUsed, [&](int k) {
    if (amb) amb[k] = amb_;
    if (dif) dif[k] = dif_;
    if (spec) spec[k] = spec_;
    if (pw) pw[k] = pow_;
  });
}

void Cuda2DTransferFunction::Modified() { ++mtime_; }

void Cuda2DTransferFunction::AddFunctionObject(std::shared_ptr<FunctionPolygon> p) {
  objs_.push_back(std::move(p));
  Modified();
}

void Cuda2DTransferFunction::RemoveFunctionObject(const std::shared_ptr<FunctionPolygon>& p) {
  auto it = std::find(objs_.begin(), objs_.end(), p);
  if (it != objs_.end()) {
    objs_.erase(it);
    Modified();
  }
}

std::shared_ptr<FunctionPolygon> Cuda2DTransferFunction::GetFunctionObject(size_t i) const {
  return i < objs_.size() ? objs_[i] : nullptr;
}

double Cuda2DTransferFunction::getMinIntensity() const {
  double m = objs_.empty() ? 0.0 : std::numeric_limits<double>::max();
  for (const auto& o : objs_) m = std::min(m, o->getMinIntensity());
  return m;
}
double Cuda2DTransferFunction::getMaxIntensity() const {
  double m = objs_.empty() ? 0.0 : -std::numeric_limits<double>::max();
  for (const auto& o : objs_) m = std::max(m, o->getMaxIntensity());
  return m;
}
double Cuda2DTransferFunction::getMinGradient() const {
  double m = objs_.empty() ? 0.0 : std::numeric_limits<double>::max();
  for (const auto& o : objs_) m = std::min(m, o->getMinGradient());
  return m;
}
double Cuda2DTransferFunction::getMaxGradient() const {
  double m = objs_.empty() ? 0.0 : -std::numeric_limits<double>::max();
  for (const auto& o : objs_) m = std::max(m, o->getMaxGradient());
  return m;
}

void Cuda2DTransferFunction::GetTransferTable(float* r, float* g, float* b, float* a, int sizeI, int sizeG,
                                              float lowI, float highI, float offI, float lowG, float highG,
                                              float offG, int logUsed) const {
  const size_t n = (size_t)sizeI * sizeG;
  for (float* t : {r, g, b, a})
    if (t) std::fill(t, t + n, 0.0f);
  for (const auto& o : objs_)
    o->PopulatePortionOfTransferTable(sizeI, sizeG, lowI, highI, offI, lowG, highG, offG, r, g, b, a, logUsed);
}

void Cuda2DTransferFunction::GetShadingTable(float* amb, float* dif, float* spec, float* pw, int sizeI, int sizeG,
                                             float lowI, float highI, float offI, float lowG, float highG,
                                             float offG, int logUsed) const {
  const size_t n = (size_t)sizeI * sizeG;
  for (float* t : {amb, dif, spec, pw})
    if (t) std::fill(t, t + n, 0.0f);
  for (const auto& o : objs_)
    o->PopulatePortionOfShadingTable(sizeI, sizeG, lowI, highI, offI, lowG, highG, offG, amb, dif, spec, pw,
                                     logUsed);
}

bool ReadFunctionPolygons(const std::string& path, std::vector<std::shared_ptr<FunctionPolygon>>& out) {
  out.clear();
  std::ifstream f(path);
  if (!f) return false;
  int n = 0;
  if (!(f >> n) || n < 0) return false;
  for (int k = 0; k < n; ++k) {
    auto p = std::make_shared<FunctionPolygon>();
    float r, g, b, a, amb, dif, spec, pw;
    short id;
    int nv = 0;
    if (!(f >> r >> g >> b >> a >> amb >> dif >> spec >> pw >> id >> nv) || nv < 0) return false;
    p->SetColour(r, g, b);
    p->SetOpacity(a);
    p->SetAmbient(amb);
    p->SetDiffuse(dif);
    p->SetSpecular(spec);
    p->SetSpecularPower(pw);
    p->SetIdentifier(id);
    for (int v = 0; v < nv; ++v) {
      float i, gr;
      if (!(f >> i >> gr)) return false;
      p->AddVertex(i, gr);
    }
    out.push_back(p);
  }
  return true;
}

bool WriteFunctionPolygons(const std::string& path, const std::vector<std::shared_ptr<FunctionPolygon>>& in) {
  std::ofstream f(path);
  if (!f) return false;
  f.precision(9);
  f << in.size() << "\n";
  for (const auto& p : in) {
    f << p->GetRedColourValue() << " " << p->GetGreenColourValue() << " " << p->GetBlueColourValue() << " "
      << p->GetOpacity() << "\n";
    f << p->GetAmbient() << " " << p->GetDiffuse() << " " << p->GetSpecular() << " " << p->GetSpecularPower() << "\n";
    f << p->GetIdentifier() << "\n" << p->GetNumberOfVertices() << "\n";
    for (size_t v = 0; v < p->GetNumberOfVertices(); ++v)
      f << p->GetVertexIntensity(v) << " " << p->GetVertexGradient(v) << "\n";
  }
  return (bool)f;
}

// ---- CudaObject --------------------------------------------------------------------
CudaObject::CudaObject(int d) : device_(d) {}

CudaObject::~CudaObject() {
  if (ctx_) rvk_ctx_destroy(ctx_);
}

bool CudaObject::EnsureContext() {
  if (ctx_) return true;
  if (rvk_ctx_create(device_, &ctx_) != RVK_OK) {
    ctx_ = nullptr;
    ctx_error_ = rvk_last_error_string();
    return false;
  }
  return true;
}

void CudaObject::SetDevice(int d, bool withData) {
  if (d == device_ && ctx_) return;
  Deinitialize(withData);
  if (ctx_) rvk_ctx_destroy(ctx_);
  ctx_ = nullptr;
  device_ = d;
  Reinitialize(withData);
}

void CudaObject::ReserveGPU() { EnsureContext(); }

void CudaObject::CallSyncThreads() {
  if (ctx_) rvk_ctx_sync(ctx_);
}

void* CudaObject::GetStream() {
  void* s = nullptr;
  if (EnsureContext()) rvk_ctx_stream(ctx_, &s);
  return s;
}

// ---- CudaVolumeMapper ------------------------------------------------------------------
CudaVolumeMapper::CudaVolumeMapper(float default_term) : CudaObject(0) {
  march_.step_scale = 1.0f;
  march_.term_T = default_term;
  march_.jitter = 1;
  march_.skip_rule = 1;
  march_.precise = 0;
  // V/vtkCudaRendererInformationHandler.cxx:44-45
  SetCelShadingConstants(0.70f, 0.002f, 0.005f);
  SetDistanceShadingConstants(0.25f, 0.45f, 0.6f);
  inputs_.resize(1);
  has_input_.assign(1, false);
}

CudaVolumeMapper::~CudaVolumeMapper() { Deinitialize(false); }

bool CudaVolumeMapper::Fail(const std::string& what) {
  error_ = true;
  last_error_ = what;
  return false;
}

bool CudaVolumeMapper::Check(int rc, const char* what) {
  if (rc == RVK_OK) return true;
  return Fail(std::string(what) + ": " + rvk_last_error_string());
}

void CudaVolumeMapper::DestroyVolumes() {
  if (vol_) rvk_volume_destroy(vol_);
  vol_ = nullptr;
}

void CudaVolumeMapper::Deinitialize(bool) {
  if (!ctx_) return;
  DestroyVolumes();
  if (d_out_) rvk_device_free(ctx_, d_out_);
  for (auto& a : d_aux_) {
    if (a) rvk_device_free(ctx_, a);
    a = nullptr;
  }
  if (d_zbuf_) rvk_device_free(ctx_, d_zbuf_);
  d_out_ = nullptr;
  d_zbuf_ = nullptr;
  d_out_bytes_ = d_aux_pixels_ = d_zbuf_n_ = 0;
}

// V/vtkCudaVolumeMapper.cxx:140-157: after a device move the cached inputs are
// uploaded again when withData is set.
void CudaVolumeMapper::Reinitialize(bool withData) {
  if (!withData) {
    std::fill(has_input_.begin(), has_input_.end(), false);
    return;
  }
  for (int f = 0; f < n_frames_; ++f)
    if (has_input_[f] && !UploadFrame(f)) return;
}

void CudaVolumeMapper::SetNumberOfFrames(int n) {
  if (n < 1 || n > RVK_MAX_FRAMES) {
    Fail("number of frames must be within 1.." + std::to_string(RVK_MAX_FRAMES));
    return;
  }
  if (n == n_frames_) return;
  DestroyVolumes();
  inputs_.resize(n);
  has_input_.resize(n, false);
  n_frames_ = n;
  frame_ = std::min(frame_, n - 1);
  for (int f = 0; f < n; ++f)
    if (has_input_[f] && !UploadFrame(f)) return;
}

bool CudaVolumeMapper::FitVolume(rvk_volume*& v, const ImageData& im, int f) {
  if (v) {
    int32_t d[3];
    double sp[3];
    int dt = 0, nf = 0;
    rvk_volume_dims(v, d, sp, &dt, &nf);
    if (d[0] != im.dims[0] || d[1] != im.dims[1] || d[2] != im.dims[2] || dt != im.scalar_type || nf != n_frames_ ||
        sp[0] != im.spacing[0] || sp[1] != im.spacing[1] || sp[2] != im.spacing[2]) {
      rvk_volume_destroy(v);
      v = nullptr;
      for (int g = 0; g < n_frames_; ++g)
        if (g != f) has_input_[g] = false;
    }
  }
  if (!v) {
    const int32_t d[3] = {im.dims[0], im.dims[1], im.dims[2]};
    return Check(rvk_volume_create(ctx_, d, im.spacing, im.scalar_type, n_frames_, &v), "volume create");
  }
  return true;
}

bool CudaVolumeMapper::UploadFrame(int f) {
  const ImageData& im = inputs_[f];
  if (im.components != 1) return Fail("input must have 1 scalar component");
  if (!EnsureContext()) return Fail(ctx_error_);
  if (!FitVolume(vol_, im, f)) return false;
  return Check(rvk_volume_set_frame(vol_, f, im.scalars, 0), "volume upload");
}

void CudaVolumeMapper::SetInputData(const ImageData& image, int frame) {
  if (frame < 0 || frame >= n_frames_) {
    Fail("frame " + std::to_string(frame) + " out of range");
    return;
  }
  if (!image.scalars || dtype_bytes(image.scalar_type) == 0) {
    Fail("input image has no scalars or an unsupported scalar type");
    return;
  }
  inputs_[frame] = image;
  has_input_[frame] = true;
  // the reference stores CanRender = !success on every input change
  // (V/vtkCuda1DVolumeMapper.cxx:178)
  error_ = false;
  last_error_.clear();
  if (!UploadFrame(frame)) has_input_[frame] = false;
}

const ImageData* CudaVolumeMapper::GetInput(int frame) const {
  if (frame < 0 || frame >= n_frames_ || !has_input_[frame]) return nullptr;
  return &inputs_[frame];
}

void CudaVolumeMapper::ClearInput() {
  std::fill(has_input_.begin(), has_input_.end(), false);
  DestroyVolumes();
}

void CudaVolumeMapper::ChangeFrame(int f) {
  if (f >= 0 && f < n_frames_) frame_ = f;
}

void CudaVolumeMapper::AdvanceFrame() { ChangeFrame((frame_ + 1) % n_frames_); }

void CudaVolumeMapper::SetRenderOutputScaleFactor(float s) {
  if (s >= 1.0f) scale_factor_ = s;
}

// V/vtkCudaRendererInformationHandler.cxx:79-99 (darkness outside [0,1] ignored)
void CudaVolumeMapper::SetCelShadingConstants(float darkness, float a, float b) {
  if (darkness >= 0.0f && darkness <= 1.0f) {
    shade_.cela = a;
    shade_.celr = darkness;
    shade_.celc = (float)(1.0 / (b - a));
  }
}

void CudaVolumeMapper::SetDistanceShadingConstants(float darkness, float a, float b) {
  if (darkness >= 0.0f && darkness <= 1.0f) {
    shade_.disa = a;
    shade_.disr = darkness;
    shade_.disc = (float)(1.0 / (b - a));
  }
}

void CudaVolumeMapper::SetTerminationThreshold(float t) { march_.term_T = t; }

void CudaVolumeMapper::AddKeyholePlane(const Plane& p) {
  if (keyhole_.size() < 6) keyhole_.push_back(p);
}

void CudaVolumeMapper::RemoveKeyholePlane(size_t i) {
  if (i < keyhole_.size()) keyhole_.erase(keyhole_.begin() + i);
}

void CudaVolumeMapper::AddClippingPlane(const Plane& p) {
  if (clipping_.size() < 6) clipping_.push_back(p);
}

// V/vtkCudaVolumeMapper.cxx:496-549
void CudaVolumeMapper::SetTint(const double rgba[4]) {
  for (int i = 0; i < 4; ++i) tint_[i] = (unsigned char)std::min(255.0, std::max(0.0, 255.0 * rgba[i]));
}
void CudaVolumeMapper::SetTint(const float rgba[4]) {
  for (int i = 0; i < 4; ++i) tint_[i] = (unsigned char)std::min(255.0f, std::max(0.0f, 255.0f * rgba[i]));
}
void CudaVolumeMapper::SetTint(const unsigned char rgba[4]) { std::memcpy(tint_, rgba, 4); }
void CudaVolumeMapper::GetTint(double rgba[4]) const {
  for (int i = 0; i < 4; ++i) rgba[i] = (double)tint_[i] / 255.0;
}
void CudaVolumeMapper::GetTint(float rgba[4]) const {
  for (int i = 0; i < 4; ++i) rgba[i] = (float)tint_[i] / 255.0f;
}
void CudaVolumeMapper::GetTint(unsigned char rgba[4]) const { std::memcpy(rgba, tint_, 4); }

bool CudaVolumeMapper::EnsureOutput(size_t bytes, bool aux) {
  if (bytes > d_out_bytes_) {
    if (d_out_) rvk_device_free(ctx_, d_out_);
    d_out_ = nullptr;
    if (!Check(rvk_device_alloc(ctx_, bytes, &d_out_), "output allocation")) return false;
    d_out_bytes_ = bytes;
  }
  const size_t px = (size_t)view_.width * view_.height;
  if (aux && px > d_aux_pixels_) {
    for (auto& a : d_aux_) {
      if (a) rvk_device_free(ctx_, a);
      a = nullptr;
      if (!Check(rvk_device_alloc(ctx_, px * sizeof(float), (void**)&a), "aux allocation")) return false;
    }
    d_aux_pixels_ = px;
  }
  return true;
}

bool CudaVolumeMapper::UploadZBuffer(const Renderer& ren) {
  view_.zbuffer = nullptr;
  if (!ren.zbuffer) return true;
  const size_t n = (size_t)ren.size[0] * ren.size[1];
  if (n > d_zbuf_n_) {
    if (d_zbuf_) rvk_device_free(ctx_, d_zbuf_);
    d_zbuf_ = nullptr;
    if (!Check(rvk_device_alloc(ctx_, n * sizeof(float), (void**)&d_zbuf_), "z-buffer allocation")) return false;
    d_zbuf_n_ = n;
  }
  if (!Check(rvk_copy_to_device(ctx_, d_zbuf_, ren.zbuffer, n * sizeof(float)), "z-buffer upload")) return false;
  view_.zbuffer = d_zbuf_;
  view_.zb_width = ren.size[0];
  view_.zb_height = ren.size[1];
  return true;
}

// ComputeMatrices (V/vtkCudaVolumeMapper.cxx:349-447) with the output-size
// policy of V/vtkCudaOutputImageInformationHandler.cxx:279-301 (padding only
// on request) and FigurePlanes (V/vtkCudaRendererInformationHandler.cxx:216-260:
// planes are used only when there are exactly six).
bool CudaVolumeMapper::ComputeView(const Renderer& ren, const Volume& vol) {
  const ImageData& im = inputs_[frame_];
  int W = (int)(ren.size[0] / scale_factor_), H = (int)(ren.size[1] / scale_factor_);
  if (pad_) {
    W += (W % 16) ? 16 - (W % 16) : 0;
    H += (H % 16) ? 16 - (H % 16) : 0;
    W = std::max(W, 256);
    H = std::max(H, 256);
  }
  if (W < 1 || H < 1) return Fail("renderer size is empty");
  const rvk_camera cam = ren.camera.ToABI();
  const double* um = vol.has_user_matrix ? vol.user_matrix : nullptr;
  double v2vox[16], vox2v[16];
  if (!Check(rvk_compute_matrices(&cam, um, im.origin, im.spacing, (double)ren.size[0] / (double)ren.size[1], v2vox,
                                  vox2v),
             "camera matrices"))
    return false;
  std::memset(&view_, 0, sizeof(view_));
  rvk_prebake_view_to_voxels(v2vox, view_.view_to_voxels);
  view_.width = W;
  view_.height = H;
  view_.flipped = flipped_;
  // voxels -> world and its inverse for the planes
  double vox[16] = {im.spacing[0], 0, 0, im.origin[0], 0, im.spacing[1], 0, im.origin[1],
                    0, 0, im.spacing[2], im.origin[2], 0, 0, 0, 1};
  double v2w[16], w2v[16];
  if (um) rvk_mat4_mul(um, vox, v2w);
  else std::memcpy(v2w, vox, sizeof(vox));
  if (rvk_mat4_invert(v2w, w2v) != RVK_OK) return Fail("volume matrix is singular");
  auto figure = [&](const std::vector<Plane>& planes, int32_t* count, float* out) {
    *count = 0;
    if (planes.size() != 6) return;
    *count = 6;
    for (int i = 0; i < 6; ++i) {
      const double* n = planes[i].normal;
      const double* o = planes[i].origin;
      for (int c = 0; c < 3; ++c) out[4 * i + c] = (float)(n[0] * v2w[c] + n[1] * v2w[4 + c] + n[2] * v2w[8 + c]);
      double vo[4];
      for (int r = 0; r < 4; ++r) vo[r] = o[0] * w2v[4 * r] + o[1] * w2v[4 * r + 1] + o[2] * w2v[4 * r + 2] + w2v[4 * r + 3];
      if (vo[3] != 1.0)
        for (int r = 0; r < 3; ++r) vo[r] /= vo[3];
      out[4 * i + 3] = (float)-((double)out[4 * i] * vo[0] + (double)out[4 * i + 1] * vo[1] +
                                (double)out[4 * i + 2] * vo[2]);
    }
  };
  figure(clipping_, &view_.n_clip_planes, view_.clip_planes);
  figure(keyhole_, &view_.n_keyhole_planes, view_.keyhole_planes);
  return UploadZBuffer(ren);
}

void CudaVolumeMapper::Render(const Renderer& ren, const Volume& vol) {
  // V/vtkCudaVolumeMapper.cxx:301-340: an earlier error blocks rendering
  if (error_) return;
  if (frame_ < 0 || frame_ >= n_frames_ || !has_input_[frame_] || !vol_) {
    Fail("Render called without input data for the current frame");
    return;
  }
  if (!EnsureContext()) {
    Fail(ctx_error_);
    return;
  }
  if (!ComputeView(ren, vol)) return;
  InternalRender(ren, vol);
}

bool CudaVolumeMapper::PostShade(bool inex) {
  // with zero darkness both shades leave the image as it is
  if (!post_shade_ || (shade_.celr == 0.0f && shade_.disr == 0.0f)) return true;
  auto fn = inex ? rvk_post_shade_inex : rvk_post_shade;
  return Check(fn(ctx_, view_.width, view_.height, &shade_, d_aux_[0], d_aux_[1], d_aux_[2], (uint8_t*)d_out_),
               "post shading");
}

bool CudaVolumeMapper::FinishRGBA() {
  const size_t px = (size_t)view_.width * view_.height;
  output_.width = view_.width;
  output_.height = view_.height;
  output_.values.clear();
  output_.rgba.resize(px * 4);
  return Check(rvk_copy_to_host(ctx_, output_.rgba.data(), d_out_, px * 4), "output copy");
}

// ---- 1D ---------------------------------------------------------------------------------
Cuda1DVolumeMapper::Cuda1DVolumeMapper() : CudaVolumeMapper(1.0f / 64.0f) {}
Cuda1DVolumeMapper::~Cuda1DVolumeMapper() = default;

// V/vtkCuda1DTransferFunctionInformationHandler.cxx:164-273 and
// V/vtkCuda1DVolumeMapper.cxx:195-214
void Cuda1DVolumeMapper::InternalRender(const Renderer&, const Volume& vol) {
  const VolumeProperty& pr = vol.property;
  if (pr.scalar_opacity.Empty()) {
    Fail("the volume property has no scalar opacity function");
    return;
  }
  double sr[2];
  inputs_[frame_].GetScalarRange(sr);
  double lo, hi;
  pr.scalar_opacity.GetRange(lo, hi);
  const double minI = std::max(sr[0], lo), maxI = std::min(sr[1], hi);
  const int n = tf_size_;
  std::vector<float> alpha(n), galpha(n, 1.0f), rgb(3 * n), r(n), g(n), b(n);
  pr.scalar_opacity.GetTable(minI, maxI, n, alpha.data());
  pr.color.GetTable(minI, maxI, n, rgb.data());
  for (int i = 0; i < n; ++i) {
    r[i] = rgb[3 * i];
    g[i] = rgb[3 * i + 1];
    b[i] = rgb[3 * i + 2];
  }
  rvk_tf1d_params p;
  p.intensity_low = (float)minI;
  p.intensity_multiplier = (float)(1.0 / (maxI - minI));
  p.gradient_low = 0.0f;
  p.gradient_multiplier = std::numeric_limits<float>::infinity();
  if (!pr.gradient_opacity.Empty()) {
    double glo, ghi;
    pr.gradient_opacity.GetRange(glo, ghi);
    pr.gradient_opacity.GetTable(glo, ghi, n, galpha.data());
    p.gradient_low = (float)glo;
    p.gradient_multiplier = (float)(1.0 / (ghi - glo));
  }
  p.ambient = 1.0f;
  p.diffuse = 0.0f;
  p.specular = 0.0f;
  p.specular_power = 1.0f;
  if (pr.shade) {
    p.ambient = (float)pr.ambient;
    p.diffuse = (float)pr.diffuse;
    p.specular = (float)pr.specular;
    p.specular_power = (float)pr.specular_power;
  }
  rvk_tf1d* tf = nullptr;
  if (!Check(rvk_tf1d_create(ctx_, n, alpha.data(), galpha.data(), r.data(), g.data(), b.data(), &p, &tf),
             "transfer function upload"))
    return;
  const size_t px = (size_t)view_.width * view_.height;
  bool ok = EnsureOutput(px * 4, true);
  if (ok) {
    rvk_aux_out aux{d_aux_[0], d_aux_[1], d_aux_[2]};
    ok = Check(rvk_render_1d(ctx_, vol_, frame_, tf, &view_, &march_, (uint8_t*)d_out_, &aux), "1D render");
  }
  rvk_tf1d_destroy(tf);
  if (ok) ok = PostShade();
  if (ok) FinishRGBA();
}

// ---- 2D ---------------------------------------------------------------------------------
Cuda2DVolumeMapper::Cuda2DVolumeMapper() : CudaVolumeMapper(1.0f / 32.0f) {}
Cuda2DVolumeMapper::~Cuda2DVolumeMapper() = default;

void Cuda2DVolumeMapper::SetTables(int n, const std::vector<std::vector<float>>& t, const rvk_tf2d_params& p) {
  direct_n_ = n;
  direct_ = t;
  direct_params_ = p;
}

// V/vtkCuda2DTransferFunctionInformationHandler.cxx:154-308
bool Cuda2DVolumeMapper::BuildTables(int n, std::vector<std::vector<float>>& t, std::vector<std::vector<float>>& kt,
                                     rvk_tf2d_params& p) {
  const bool dual = kfn_ && kfn_->GetNumberOfFunctionObjects() > 0;
  double fr[4] = {fn_->getMinIntensity(), fn_->getMaxIntensity(), fn_->getMinGradient(), fn_->getMaxGradient()};
  if (dual) {
    fr[0] = std::min(fr[0], kfn_->getMinIntensity());
    fr[1] = std::max(fr[1], kfn_->getMaxIntensity());
    fr[2] = std::min(fr[2], kfn_->getMinGradient());
    fr[3] = std::max(fr[3], kfn_->getMaxGradient());
  }
  const ImageData& im = inputs_[frame_];
  double sr[2];
  im.GetScalarRange(sr);
  const double minI = std::max(sr[0], fr[0]), maxI = std::min(sr[1], fr[1]);
  const double highG =
      std::fabs(sr[0] - sr[1]) / std::min(im.spacing[0], std::min(im.spacing[1], im.spacing[2]));
  double minG = std::max(0.0, fr[2]);
  double maxG = std::min(highG, fr[3]);
  const double goff = maxG * 0.8;
  maxG = (std::log(maxG * maxG + goff) - std::log(goff)) / std::log(2.0) + 1.0;
  minG = (std::log(minG * minG + goff) - std::log(goff)) / std::log(2.0);
  p.intensity_low = (float)minI;
  p.intensity_multiplier = (float)(1.0 / (maxI - minI));
  p.gradient_multiplier = (float)(1.0 / (maxG - minG));
  p.gradient_offset = (float)goff;
  p.gradient_low = (float)(-minG - std::log(goff) / std::log(2.0));
  const size_t nn = (size_t)n * n;
  t.assign(8, std::vector<float>(nn));
  fn_->GetTransferTable(t[1].data(), t[2].data(), t[3].data(), t[0].data(), n, n, (float)minI, (float)maxI, 0.0f,
                        (float)minG, (float)maxG, (float)goff, 2);
  fn_->GetShadingTable(t[4].data(), t[5].data(), t[6].data(), t[7].data(), n, n, (float)minI, (float)maxI, 0.0f,
                       (float)minG, (float)maxG, (float)goff, 2);
  kt.clear();
  if (dual) {
    kt.assign(8, std::vector<float>(nn));
    kfn_->GetTransferTable(kt[1].data(), kt[2].data(), kt[3].data(), kt[0].data(), n, n, (float)minI, (float)maxI,
                           0.0f, (float)minG, (float)maxG, (float)goff, 2);
    kfn_->GetShadingTable(kt[4].data(), kt[5].data(), kt[6].data(), kt[7].data(), n, n, (float)minI, (float)maxI,
                          0.0f, (float)minG, (float)maxG, (float)goff, 2);
  }
  return true;
}

void Cuda2DVolumeMapper::InternalRender(const Renderer&, const Volume&) {
  std::vector<std::vector<float>> t, kt;
  rvk_tf2d_params p;
  int n = tf_size_;
  if (direct_n_ > 0) {
    n = direct_n_;
    t = direct_;
    p = direct_params_;
    if (t.size() != 8) {
      Fail("SetTables needs 8 tables");
      return;
    }
  } else {
    if (!fn_ || fn_->GetNumberOfFunctionObjects() == 0) {
      Fail("no 2D transfer function set");
      return;
    }
    BuildTables(n, t, kt, p);
  }
  const float* tp[8];
  const float* kp[8];
  for (int i = 0; i < 8; ++i) {
    tp[i] = t[i].data();
    kp[i] = kt.empty() ? nullptr : kt[i].data();
  }
  rvk_tf2d* tf = nullptr;
  if (!Check(rvk_tf2d_create(ctx_, n, tp, kt.empty() ? nullptr : kp, &p, &tf), "2D transfer function upload"))
    return;
  const bool gm = gm_max_ > 0.0f;
  bool ok = true;
  if (gm && gm_frame_ != frame_) {
    ok = Check(rvk_volume_compute_gradmag(vol_, frame_, gm_max_), "gradient magnitude volume");
    gm_frame_ = ok ? frame_ : -1;
  }
  const size_t px = (size_t)view_.width * view_.height;
  if (ok) ok = EnsureOutput(px * 4, true);
  if (ok) {
    rvk_aux_out aux{d_aux_[0], d_aux_[1], d_aux_[2]};
    ok = Check(rvk_render_2d(ctx_, vol_, frame_, tf, &view_, &march_, gm, (uint8_t*)d_out_, &aux), "2D render");
  }
  rvk_tf2d_destroy(tf);
  if (ok) ok = PostShade();
  if (ok) FinishRGBA();
}

// ---- 2D with in/ex logic ----------------------------------------------------------------
Cuda2DInExLogicVolumeMapper::Cuda2DInExLogicVolumeMapper() : CudaVolumeMapper(1.0f / 32.0f) {}
Cuda2DInExLogicVolumeMapper::~Cuda2DInExLogicVolumeMapper() = default;

// V/vtkCuda2DInExLogicTransferFunctionInformationHandler.cxx:137-217
void Cuda2DInExLogicVolumeMapper::InternalRender(const Renderer&, const Volume&) {
  if (!fn_ || !inex_fn_) {
    Fail("the visualization and in/ex logic functions must both be set");
    return;
  }
  const ImageData& im = inputs_[frame_];
  double sr[2];
  im.GetScalarRange(sr);
  const double minI = std::max(sr[0], (double)fn_->getMinIntensity());
  const double maxI = std::min(sr[1], (double)fn_->getMaxIntensity());
  const double highG =
      std::fabs(sr[0] - sr[1]) / std::min(im.spacing[0], std::min(im.spacing[1], im.spacing[2]));
  double minG = std::max({0.0, (double)fn_->getMinGradient(), (double)inex_fn_->getMinGradient()});
  double maxG = std::min({highG, (double)fn_->getMaxGradient(), (double)inex_fn_->getMaxGradient()});
  const double goff = maxG * 0.8;
  maxG = (std::log(maxG * maxG + goff) - std::log(goff)) / std::log(2.0) + 1.0;
  minG = (std::log(minG * minG + goff) - std::log(goff)) / std::log(2.0);
  rvk_tf2d_params p;
  p.intensity_low = (float)minI;
  p.intensity_multiplier = (float)(1.0 / (maxI - minI));
  p.gradient_multiplier = (float)(1.0 / (maxG - minG));
  p.gradient_offset = (float)goff;
  p.gradient_low = (float)(-minG - std::log(goff) / std::log(2.0));
  const int n = tf_size_;
  const size_t nn = (size_t)n * n;
  std::vector<std::vector<float>> t(8, std::vector<float>(nn));
  std::vector<float> inex(nn);
  const float a[6] = {(float)minI, (float)maxI, 0.0f, (float)minG, (float)maxG, (float)goff};
  fn_->GetTransferTable(t[1].data(), t[2].data(), t[3].data(), t[0].data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5],
                        2);
  fn_->GetShadingTable(t[4].data(), t[5].data(), t[6].data(), t[7].data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5],
                       2);
  inex_fn_->GetTransferTable(nullptr, nullptr, nullptr, inex.data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5], 2);
  const float* tp[8];
  for (int i = 0; i < 8; ++i) tp[i] = t[i].data();
  rvk_tf2d* tf = nullptr;
  if (!Check(rvk_tf2d_create(ctx_, n, tp, nullptr, &p, &tf), "2D transfer function upload")) return;
  bool ok = Check(rvk_tf2d_set_inex(tf, inex.data()), "in/ex table upload");
  const size_t px = (size_t)view_.width * view_.height;
  if (ok) ok = EnsureOutput(px * 4, true);
  if (ok) {
    rvk_aux_out aux{d_aux_[0], d_aux_[1], d_aux_[2]};
    ok = Check(rvk_render_2d_inex(ctx_, vol_, frame_, tf, &view_, &march_, black_, (uint8_t*)d_out_, &aux),
               "in/ex render");
  }
  rvk_tf2d_destroy(tf);
  if (ok) ok = PostShade(true);
  if (ok) FinishRGBA();
}

// ---- dual image ----------------------------------------------------------------------------
CudaDualImageVolumeMapper::CudaDualImageVolumeMapper() : CudaVolumeMapper(1.0f / 32.0f) {}
CudaDualImageVolumeMapper::~CudaDualImageVolumeMapper() {
  if (vol2_) rvk_volume_destroy(vol2_);
  vol2_ = nullptr;
}

void CudaDualImageVolumeMapper::DestroyVolumes() {
  CudaVolumeMapper::DestroyVolumes();
  if (vol2_) rvk_volume_destroy(vol2_);
  vol2_ = nullptr;
}

// V/vtkCudaDualImageVolumeMapper.cxx:76-160: a 2-component input, kept as one
// device volume per component.
bool CudaDualImageVolumeMapper::UploadFrame(int f) {
  const ImageData& im = inputs_[f];
  if (im.components != 2) return Fail("Input must have 2 components.");
  if (!EnsureContext()) return Fail(ctx_error_);
  if (!FitVolume(vol_, im, f) || !FitVolume(vol2_, im, f)) return false;
  const size_t n = (size_t)im.dims[0] * im.dims[1] * im.dims[2], b = dtype_bytes(im.scalar_type);
  std::vector<uint8_t> c1(n * b), c2(n * b);
  const uint8_t* src = (const uint8_t*)im.scalars;
  for (size_t i = 0; i < n; ++i) {
    std::memcpy(&c1[i * b], src + 2 * i * b, b);
    std::memcpy(&c2[i * b], src + (2 * i + 1) * b, b);
  }
  return Check(rvk_volume_set_frame(vol_, f, c1.data(), 0), "volume upload") &&
         Check(rvk_volume_set_frame(vol2_, f, c2.data(), 0), "volume upload");
}

// V/vtkCudaDualImageTransferFunctionInformationHandler.cxx:143-290
void CudaDualImageVolumeMapper::InternalRender(const Renderer&, const Volume&) {
  if (!fn_ || fn_->GetNumberOfFunctionObjects() == 0) {
    Fail("no 2D transfer function set");
    return;
  }
  if (!vol2_) {
    Fail("Render called without input data for the current frame");
    return;
  }
  const bool key = kfn_ && kfn_->GetNumberOfFunctionObjects() > 0;
  const ImageData& im = inputs_[frame_];
  double sr[4];
  im.GetScalarRange(sr, 0);
  im.GetScalarRange(sr + 2, 1);
  double fr[4] = {fn_->getMinIntensity(), fn_->getMaxIntensity(), fn_->getMinGradient(), fn_->getMaxGradient()};
  if (key) {
    fr[0] = std::min(fr[0], (double)kfn_->getMinIntensity());
    fr[1] = std::max(fr[1], (double)kfn_->getMaxIntensity());
    fr[2] = std::min(fr[2], (double)kfn_->getMinGradient());
    fr[3] = std::max(fr[3], (double)kfn_->getMaxGradient());
  }
  const double lo1 = std::max(sr[0], fr[0]), hi1 = std::min(sr[1], fr[1]);
  const double lo2 = std::max(sr[2], fr[2]), hi2 = std::min(sr[3], fr[3]);
  const rvk_dual_params dp{(float)lo1, (float)(1.0 / (hi1 - lo1)), (float)lo2, (float)(1.0 / (hi2 - lo2))};
  const int n = tf_size_;
  const size_t nn = (size_t)n * n;
  const float a[6] = {(float)lo1, (float)hi1, 0.0f, (float)lo2, (float)hi2, 0.0f};
  std::vector<std::vector<float>> t(8, std::vector<float>(nn)), kt;
  fn_->GetTransferTable(t[1].data(), t[2].data(), t[3].data(), t[0].data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5],
                        0);
  fn_->GetShadingTable(t[4].data(), t[5].data(), t[6].data(), t[7].data(), n, n, a[0], a[1], a[2], a[3], a[4], a[5],
                       0);
  if (key) {
    kt.assign(8, std::vector<float>(nn));
    kfn_->GetTransferTable(kt[1].data(), kt[2].data(), kt[3].data(), kt[0].data(), n, n, a[0], a[1], a[2], a[3], a[4],
                           a[5], 0);
    kfn_->GetShadingTable(kt[4].data(), kt[5].data(), kt[6].data(), kt[7].data(), n, n, a[0], a[1], a[2], a[3], a[4],
                          a[5], 0);
  }
  const float* tp[8];
  const float* kp[8];
  for (int i = 0; i < 8; ++i) {
    tp[i] = t[i].data();
    kp[i] = key ? kt[i].data() : nullptr;
  }
  const rvk_tf2d_params unused{};  // the dual caster takes its scalars from dp
  rvk_tf2d* tf = nullptr;
  if (!Check(rvk_tf2d_create(ctx_, n, tp, key ? kp : nullptr, &unused, &tf), "2D transfer function upload")) return;
  const size_t px = (size_t)view_.width * view_.height;
  bool ok = EnsureOutput(px * 4, true);
  if (ok) {
    rvk_aux_out aux{d_aux_[0], d_aux_[1], d_aux_[2]};
    ok = Check(rvk_render_dual(ctx_, vol_, vol2_, frame_, tf, &dp, &view_, &march_, (uint8_t*)d_out_, &aux),
               "dual-image render");
  }
  rvk_tf2d_destroy(tf);
  if (ok) ok = PostShade();
  if (ok) FinishRGBA();
}

// ---- DRR --------------------------------------------------------------------------------
CudaDRRImageVolumeMapper::CudaDRRImageVolumeMapper() : CudaVolumeMapper(0.0f) {
  // V/vtkCudaDRRImageVolumeMapper.cxx:21-23
  drr_.ct_intercept = 0.022f;
  drr_.ct_slope = 0.000022f;
  drr_.ct_offset = 1.0f;
  drr_.output = RVK_DRR_RGBA8_REF;
  drr_.ref_termination = 1;
}
CudaDRRImageVolumeMapper::~CudaDRRImageVolumeMapper() = default;

void CudaDRRImageVolumeMapper::InternalRender(const Renderer&, const Volume&) {
  std::memcpy(drr_.tint, tint_, 4);
  const size_t px = (size_t)view_.width * view_.height;
  const bool rgba = drr_.output == RVK_DRR_RGBA8_REF;
  if (!EnsureOutput(px * 4, false)) return;
  if (!Check(rvk_render_drr(ctx_, vol_, frame_, &view_, &march_, &drr_, d_out_), "DRR render")) return;
  output_.width = view_.width;
  output_.height = view_.height;
  if (rgba) {
    output_.values.clear();
    output_.rgba.resize(px * 4);
    Check(rvk_copy_to_host(ctx_, output_.rgba.data(), d_out_, px * 4), "output copy");
  } else {
    output_.rgba.clear();
    output_.values.resize(px);
    Check(rvk_copy_to_host(ctx_, output_.values.data(), d_out_, px * 4), "output copy");
  }
}

}  // namespace rvk
