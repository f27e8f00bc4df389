// rvk_mapper.h -- VTK-free C++ volume mappers over the librvkhip C-ABI.
//
// Same method surface as RobartsVTK's CUDA mappers (V/ =
// Visualization/vtkCudaVisualization, C/ = Common/vtkCudaCommon):
//   CudaObject                <- C/CudaObject.h:44-51
//   CudaVolumeMapper          <- V/vtkCudaVolumeMapper.h:73-195
//   Cuda1DVolumeMapper        <- V/vtkCuda1DVolumeMapper.h:30-64
//   Cuda2DVolumeMapper        <- V/vtkCuda2DVolumeMapper.h:30-92
//   CudaDRRImageVolumeMapper  <- V/vtkCudaDRRImageVolumeMapper.h:30-75
//   Cuda2DInExLogicVolumeMapper <- V/vtkCuda2DInExLogicVolumeMapper.h:30-110
//   CudaDualImageVolumeMapper <- V/vtkCudaDualImageVolumeMapper.h:30-95
//   Cuda2DTransferFunction    <- C/vtkCuda2DTransferFunction.h
//   FunctionPolygon(+Reader/Writer, the .2tf format) <- C/vtkCudaFunctionPolygon*.{h,cxx}
// VTK is not available, so the VTK objects the mappers read (vtkImageData,
// vtkRenderer + vtkCamera, vtkVolume + vtkVolumeProperty, vtkPiecewiseFunction,
// vtkColorTransferFunction, vtkPlane) are small value types here.
//
// Differences on purpose: Render() leaves its image in host memory
// (GetOutput()) instead of drawing through OpenGL; the output is not padded to
// multiples of 16 unless SetPadResolution(true); errors latch (HasError)
// instead of the reference's inverted CanRender flag; there is no
// process-wide mutex (no global device state).
#ifndef RVK_MAPPER_H
#define RVK_MAPPER_H

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "rvk.h"
#include "rvk_geometry.h"

namespace rvk {

// ---- VTK stand-ins ------------------------------------------------------------
struct ImageData {  // vtkImageData (point scalars)
  int dims[3] = {0, 0, 0};
  double spacing[3] = {1, 1, 1};
  double origin[3] = {0, 0, 0};
  int scalar_type = RVK_U8;  // rvk_dtype
  int components = 1;        // interleaved per voxel; the dual-image mapper reads 2, the others 1
  const void* scalars = nullptr;  // borrowed, x fastest
  void GetScalarRange(double range[2], int component = 0) const;
};

struct Plane {  // vtkPlane
  double origin[3] = {0, 0, 0};
  double normal[3] = {0, 0, 1};
};

struct Camera {  // vtkCamera
  double position[3] = {0, 0, 1};
  double focal_point[3] = {0, 0, 0};
  double view_up[3] = {0, 1, 0};
  double view_angle = 30.0;
  double clipping_range[2] = {0.01, 1000.01};
  bool parallel_projection = false;
  double parallel_scale = 1.0;
  rvk_camera ToABI() const;
};

struct Renderer {  // vtkRenderer + its window's z-buffer
  int size[2] = {300, 300};
  Camera camera;
  const float* zbuffer = nullptr;  // optional host depth buffer, size[0] x size[1], row major
};

class PiecewiseFunction {  // vtkPiecewiseFunction (clamped)
 public:
  void AddPoint(double x, double y);
  void RemoveAllPoints();
  void GetRange(double& lo, double& hi) const;
  bool Empty() const { return pts_.empty(); }
  double GetValue(double x) const;
  // n samples at lo + i (hi - lo)/(n - 1)
  void GetTable(double lo, double hi, int n, float* out) const;

 private:
  std::vector<std::array<double, 2>> pts_;  // sorted by x
};

class ColorTransferFunction {  // vtkColorTransferFunction (RGB, clamped)
 public:
  void AddRGBPoint(double x, double r, double g, double b);
  void RemoveAllPoints();
  void GetRange(double& lo, double& hi) const;
  void GetTable(double lo, double hi, int n, float* rgb) const;  // n x 3

 private:
  PiecewiseFunction r_, g_, b_;
};

struct VolumeProperty {  // vtkVolumeProperty
  PiecewiseFunction scalar_opacity;
  PiecewiseFunction gradient_opacity;  // empty = none
  ColorTransferFunction color;
  bool shade = false;
  double ambient = 0.1, diffuse = 0.7, specular = 0.2, specular_power = 10.0;
};

struct Volume {  // vtkVolume
  VolumeProperty property;
  bool has_user_matrix = false;
  double user_matrix[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

struct OutputImage {  // headless GetOutput(): RGBA8 or float
  int width = 0, height = 0;
  std::vector<uint8_t> rgba;  // width*height*4 for the compositing casters / RGBA8 DRR
  std::vector<float> values;  // width*height for the float DRR modes
};

// ---- 2D transfer functions -----------------------------------------------------
// A polygon in (intensity, gradient magnitude) space with colour, opacity and
// shading; rasterised into tables by an even-odd point-in-polygon test
// (C/vtkCudaFunctionPolygon.cxx:179-435).
class FunctionPolygon {
 public:
  void AddVertex(float intensity, float gradient);
  void ModifyVertex(size_t i, float intensity, float gradient);
  void RemoveVertex(size_t i);
  size_t GetNumberOfVertices() const { return contour_.size(); }
  float GetVertexIntensity(size_t i) const { return contour_[i][0]; }
  float GetVertexGradient(size_t i) const { return contour_[i][1]; }
  void SetColour(float r, float g, float b) { r_ = r; g_ = g; b_ = b; }
  void SetOpacity(float a) { a_ = a; }
  void SetAmbient(float v) { amb_ = v; }
  void SetDiffuse(float v) { dif_ = v; }
  void SetSpecular(float v) { spec_ = v; }
  void SetSpecularPower(float v) { pow_ = v; }
  void SetIdentifier(short id) { id_ = id; }
  float GetRedColourValue() const { return r_; }
  float GetGreenColourValue() const { return g_; }
  float GetBlueColourValue() const { return b_; }
  float GetOpacity() const { return a_; }
  float GetAmbient() const { return amb_; }
  float GetDiffuse() const { return dif_; }
  float GetSpecular() const { return spec_; }
  float GetSpecularPower() const { return pow_; }
  short GetIdentifier() const { return id_; }
  double getMinIntensity() const;
  double getMaxIntensity() const;
  double getMinGradient() const;
  double getMaxGradient() const;
  bool PointInPolygon(float x, float y) const;
  // logUsed bit 1: intensity axis log-scaled; bit 2: gradient axis log-scaled
  void PopulatePortionOfTransferTable(int sizeI, int sizeG, float lowI, float highI, float offI, float lowG,
                                      float highG, float offG, float* r, float* g, float* b, float* a,
                                      int logUsed) const;
  void PopulatePortionOfShadingTable(int sizeI, int sizeG, float lowI, float highI, float offI, float lowG,
                                     float highG, float offG, float* amb, float* dif, float* spec, float* pw,
                                     int logUsed) const;

 private:
  template <typename F>
  void Rasterise(int sizeI, int sizeG, float lowI, float highI, float offI, float lowG, float highG, float offG,
                 int logUsed, F&& write) const;
  std::vector<std::array<float, 2>> contour_;
  float r_ = 0, g_ = 0, b_ = 0, a_ = 0, amb_ = 1, dif_ = 0, spec_ = 0, pow_ = 1;
  short id_ = 0;
};

class Cuda2DTransferFunction {  // C/vtkCuda2DTransferFunction.cxx:96-230
 public:
  void AddFunctionObject(std::shared_ptr<FunctionPolygon> p);
  void RemoveFunctionObject(const std::shared_ptr<FunctionPolygon>& p);
  size_t GetNumberOfFunctionObjects() const { return objs_.size(); }
  std::shared_ptr<FunctionPolygon> GetFunctionObject(size_t i) const;
  double getMinIntensity() const;
  double getMaxIntensity() const;
  double getMinGradient() const;
  double getMaxGradient() const;
  void GetTransferTable(float* r, float* g, float* b, float* a, int sizeI, int sizeG, float lowI, float highI,
                        float offI, float lowG, float highG, float offG, int logUsed) const;
  void GetShadingTable(float* amb, float* dif, float* spec, float* pw, int sizeI, int sizeG, float lowI, float highI,
                       float offI, float lowG, float highG, float offG, int logUsed) const;
  uint64_t GetMTime() const { return mtime_; }
  void Modified();

 private:
  std::vector<std::shared_ptr<FunctionPolygon>> objs_;
  uint64_t mtime_ = 1;
};

// The reference's plain-text .2tf format: object count, then per polygon
// "r g b a", "ambient diffuse specular power", identifier, vertex count and
// "intensity gradient" pairs (C/vtkCudaFunctionPolygonReader.cxx:75-136).
bool ReadFunctionPolygons(const std::string& path, std::vector<std::shared_ptr<FunctionPolygon>>& out);
bool WriteFunctionPolygons(const std::string& path, const std::vector<std::shared_ptr<FunctionPolygon>>& in);

// ---- device object --------------------------------------------------------------
class CudaObject {  // C/CudaObject.h:44-51
 public:
  // Move to device d; withData re-uploads the cached inputs (the reference's
  // Deinitialize / Reinitialize(withData) pair, C/CudaObject.cxx:95-156).
  void SetDevice(int d, bool withData = false);
  int GetDevice() const { return device_; }
  void ReserveGPU();
  void CallSyncThreads();
  void* GetStream();  // hipStream_t
  rvk_ctx* GetContext() { return ctx_; }
  virtual ~CudaObject();

 protected:
  explicit CudaObject(int d = 0);
  virtual void Reinitialize(bool withData = false) = 0;
  virtual void Deinitialize(bool withData = false) = 0;
  bool EnsureContext();
  int device_ = 0;
  rvk_ctx* ctx_ = nullptr;
  std::string ctx_error_;
};

// ---- mappers ----------------------------------------------------------------------
class CudaVolumeMapper : public CudaObject {  // V/vtkCudaVolumeMapper.h:73-195
 public:
  ~CudaVolumeMapper() override;
  void SetInputData(const ImageData& image) { SetInputData(image, 0); }
  void SetInputData(const ImageData& image, int frame);
  const ImageData* GetInput() const { return GetInput(0); }
  const ImageData* GetInput(int frame) const;
  void ClearInput();
  void Render(const Renderer& renderer, const Volume& volume);
  void SetRenderOutputScaleFactor(float s);
  void SetCelShadingConstants(float darkness, float a, float b);
  void SetDistanceShadingConstants(float darkness, float a, float b);
  void ChangeFrame(int frame);
  int GetCurrentFrame() const { return frame_; }
  void AdvanceFrame();
  void SetNumberOfFrames(int n);
  int GetNumberOfFrames() const { return n_frames_; }
  const OutputImage& GetOutput() const { return output_; }
  void AddKeyholePlane(const Plane& p);
  void RemoveKeyholePlane(size_t i);
  void RemoveAllKeyholePlanes() { keyhole_.clear(); }
  void AddClippingPlane(const Plane& p);  // vtkAbstractMapper3D clipping planes
  void RemoveAllClippingPlanes() { clipping_.clear(); }
  void SetImageFlipped(bool b) { flipped_ = b; }
  bool GetImageFlipped() const { return flipped_; }
  void SetTint(const double rgba[4]);
  void SetTint(const float rgba[4]);
  void SetTint(const unsigned char rgba[4]);
  void GetTint(double rgba[4]) const;
  void GetTint(float rgba[4]) const;
  void GetTint(unsigned char rgba[4]) const;
  // additions: march controls (reference defaults: 1 MinSpacing, jitter and
  // skip rule on, reference FP32 positions)
  void SetSampleDistance(float step_in_min_spacing) { march_.step_scale = step_in_min_spacing; }
  float GetSampleDistance() const { return march_.step_scale; }
  void SetTerminationThreshold(float t);
  void SetJitter(bool on) { march_.jitter = on; }
  void SetSkipRule(bool on) { march_.skip_rule = on; }
  void SetPreciseMarch(bool on) { march_.precise = on; }
  void SetPadResolution(bool on) { pad_ = on; }
  void SetPostShading(bool on) { post_shade_ = on; }
  bool HasError() const { return error_; }
  const std::string& GetLastError() const { return last_error_; }
  // The last computed ViewToVoxels view (after Render).
  const rvk_view& GetView() const { return view_; }

 protected:
  explicit CudaVolumeMapper(float default_term);
  virtual void InternalRender(const Renderer& ren, const Volume& vol) = 0;
  void Reinitialize(bool withData) override;
  void Deinitialize(bool withData) override;
  bool Fail(const std::string& what);
  bool Check(int rc, const char* what);
  virtual bool UploadFrame(int frame);
  virtual void DestroyVolumes();
  // (Re)creates *v to hold frames like im; a change of size, type or spacing
  // drops the other frames' inputs.
  bool FitVolume(rvk_volume*& v, const ImageData& im, int frame);
  bool ComputeView(const Renderer& ren, const Volume& vol);
  bool PostShade(bool inex = false);
  bool FinishRGBA();  // copies the device RGBA8 image to output_
  rvk_volume* vol_ = nullptr;
  std::vector<ImageData> inputs_;
  std::vector<bool> has_input_;
  int frame_ = 0, n_frames_ = 1;
  rvk_march_params march_{};
  rvk_view view_{};
  OutputImage output_;
  bool error_ = false;
  std::string last_error_;
  bool flipped_ = false, pad_ = false, post_shade_ = true;
  float scale_factor_ = 1.0f;
  unsigned char tint_[4] = {0, 0, 0, 0};
  rvk_shade_params shade_{};
  std::vector<Plane> keyhole_, clipping_;
  // device scratch for the output image and the shading side outputs
  void* d_out_ = nullptr;
  float* d_aux_[3] = {nullptr, nullptr, nullptr};
  size_t d_out_bytes_ = 0, d_aux_pixels_ = 0;
  float* d_zbuf_ = nullptr;
  size_t d_zbuf_n_ = 0;
  bool EnsureOutput(size_t bytes, bool aux);
  bool UploadZBuffer(const Renderer& ren);
};

class Cuda1DVolumeMapper : public CudaVolumeMapper {  // V/vtkCuda1DVolumeMapper.cxx
 public:
  Cuda1DVolumeMapper();
  ~Cuda1DVolumeMapper() override;
  void SetTransferFunctionSize(int n) { tf_size_ = n; }

 protected:
  void InternalRender(const Renderer& ren, const Volume& vol) override;
  int tf_size_ = 512;  // V/vtkCuda1DTransferFunctionInformationHandler.cxx:42
};

class Cuda2DVolumeMapper : public CudaVolumeMapper {  // V/vtkCuda2DVolumeMapper.cxx
 public:
  Cuda2DVolumeMapper();
  ~Cuda2DVolumeMapper() override;
  void SetFunction(std::shared_ptr<Cuda2DTransferFunction> f) { fn_ = std::move(f); }
  std::shared_ptr<Cuda2DTransferFunction> GetFunction() const { return fn_; }
  void SetKeyholeFunction(std::shared_ptr<Cuda2DTransferFunction> f) { kfn_ = std::move(f); }
  std::shared_ptr<Cuda2DTransferFunction> GetKeyholeFunction() const { return kfn_; }
  void SetTransferFunctionSize(int n) { tf_size_ = n; }
  // Use a precomputed u8 gradient-magnitude volume of maximum max_gradient
  // instead of six extra gathers per sample (shading then takes the
  // Phong-Lambert factor as 1); <= 0 turns it off.
  void SetGradientMagnitudeVolume(float max_gradient) { gm_max_ = max_gradient; }
  // Tables given directly (n x n each, index g*n+i, order alpha, red, green,
  // blue, ambient, diffuse, specular, specular power) instead of polygons.
  void SetTables(int n, const std::vector<std::vector<float>>& tables, const rvk_tf2d_params& params);

 protected:
  void InternalRender(const Renderer& ren, const Volume& vol) override;
  bool BuildTables(int n, std::vector<std::vector<float>>& t, std::vector<std::vector<float>>& kt,
                   rvk_tf2d_params& p);
  std::shared_ptr<Cuda2DTransferFunction> fn_, kfn_;
  int tf_size_ = 512;  // V/vtkCuda2DTransferFunctionInformationHandler.cxx:29
  float gm_max_ = 0.0f;
  int gm_frame_ = -1;
  int direct_n_ = 0;
  std::vector<std::vector<float>> direct_;
  rvk_tf2d_params direct_params_{};
};

// The 2D caster with in/ex logic.  Tables as the reference's handler builds
// them (V/vtkCuda2DInExLogicTransferFunctionInformationHandler.cxx:137-217),
// except that the gradient range is clipped by this input's estimate
// (|scalar range| / min spacing) on every build, as the 2D mapper does; the
// reference clips by the previous build's estimate (10 on the first).
class Cuda2DInExLogicVolumeMapper : public CudaVolumeMapper {
 public:
  Cuda2DInExLogicVolumeMapper();
  ~Cuda2DInExLogicVolumeMapper() override;
  void SetVisualizationFunction(std::shared_ptr<Cuda2DTransferFunction> f) { fn_ = std::move(f); }
  std::shared_ptr<Cuda2DTransferFunction> GetVisualizationFunction() const { return fn_; }
  void SetInExLogicFunction(std::shared_ptr<Cuda2DTransferFunction> f) { inex_fn_ = std::move(f); }
  std::shared_ptr<Cuda2DTransferFunction> GetInExLogicFunction() const { return inex_fn_; }
  void SetUseBlackKeyhole(bool t) { black_ = t; }
  bool GetUseBlackKeyhole() const { return black_; }
  void SetTransferFunctionSize(int n) { tf_size_ = n; }

 protected:
  void InternalRender(const Renderer& ren, const Volume& vol) override;
  std::shared_ptr<Cuda2DTransferFunction> fn_, inex_fn_;
  bool black_ = false;
  int tf_size_ = 512;  // V/vtkCuda2DInExLogicTransferFunctionInformationHandler.cxx:29
};

// The dual-image caster over a 2-component input: component 1 indexes the
// table's intensity axis, component 2 its second axis (linear, no gradient),
// shading uses component 1's gradient
// (V/vtkCudaDualImageTransferFunctionInformationHandler.cxx:143-290).
class CudaDualImageVolumeMapper : public CudaVolumeMapper {
 public:
  CudaDualImageVolumeMapper();
  ~CudaDualImageVolumeMapper() override;
  void SetFunction(std::shared_ptr<Cuda2DTransferFunction> f) { fn_ = std::move(f); }
  std::shared_ptr<Cuda2DTransferFunction> GetFunction() const { return fn_; }
  void SetKeyholeFunction(std::shared_ptr<Cuda2DTransferFunction> f) { kfn_ = std::move(f); }
  std::shared_ptr<Cuda2DTransferFunction> GetKeyholeFunction() const { return kfn_; }
  void SetTransferFunctionSize(int n) { tf_size_ = n; }

 protected:
  void InternalRender(const Renderer& ren, const Volume& vol) override;
  bool UploadFrame(int frame) override;
  void DestroyVolumes() override;
  std::shared_ptr<Cuda2DTransferFunction> fn_, kfn_;
  int tf_size_ = 512;  // V/vtkCudaDualImageTransferFunctionInformationHandler.cxx:30
  rvk_volume* vol2_ = nullptr;  // component 2 (vol_ holds component 1)
};

class CudaDRRImageVolumeMapper : public CudaVolumeMapper {  // V/vtkCudaDRRImageVolumeMapper.cxx
 public:
  CudaDRRImageVolumeMapper();
  ~CudaDRRImageVolumeMapper() override;
  void SetCTIntercept(float v) { drr_.ct_intercept = v; }
  float GetCTIntercept() const { return drr_.ct_intercept; }
  void SetCTSlope(float v) { drr_.ct_slope = v; }
  float GetCTSlope() const { return drr_.ct_slope; }
  void SetCTOffset(float v) { drr_.ct_offset = v; }
  float GetCTOffset() const { return drr_.ct_offset; }
  // RVK_DRR_RGBA8_REF (reference, default), RVK_DRR_F32_EXP_NEG_L or RVK_DRR_F32_L
  void SetOutputMode(int mode) { drr_.output = mode; }
  void SetReferenceTermination(bool on) { drr_.ref_termination = on; }

 protected:
  void InternalRender(const Renderer& ren, const Volume& vol) override;
  rvk_drr_params drr_{};
};

}  // namespace rvk

#endif
