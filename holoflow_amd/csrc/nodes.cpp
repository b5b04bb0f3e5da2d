// Holoflow node factories of the Doppler-moment path, HIP-backed.
//
// Each kind keeps the reference's settings keys, validation (std::invalid_argument from
// infer), output descriptors, in-place aliasing and create/update reuse rules, and runs
// its arithmetic through libhfx's C ABI (include/hfx.h).  Kinds (registered under the
// reference's names, manager.cc:159-232):
//   BatchQueue (async ring)     batch_queue.cc:54-306
//   Conversion                  conversion.cu:138-258     (U16 -> F32 / CF32, "Real")
//   Reshape, Slice (views)      reshape.cu, slice.cu:242-250
//   Copy, Concatenate           copy.cu:40,117-125, concatenate.cu:210,344-352
//   RFFT, FFT                   rfft.cu:148-297, fft.cu:234-386
//   FresnelDiffraction          fresnel_diffraction.cu:386-668
//   AngularSpectrum             angular_spectrum.cu:393-580
//   Filter2D, Flatfield         filter2d.cu:359-497, flatfield.cu:323-451
//   MeanAbs, Mean, Abs          mean_abs.cu:231-400, mean.cu:254-435, abs.cu:58-70
//   Arange, Multiply            arange.cu:61-200, multiply.cu:69-92
//   FFTShiftNp                  fftshift.cu:110-260
//   DopplerMoments (fused)      the whole chain graph_builder.cc:245-305, 522-566, 604-647
//   Holofile (source)           sources/holofile.cc:244-434
//   Memcpy                      memcpy.cc:105-196
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "curaii/hip.hh"
#include "hfx.h"
#include "hfx_holofile.h"
#include "hfx_nodes.h"
#include "holoflow/core/registry.hh"

namespace hfx {
void set_last_error(const std::string &msg);  // libhfx's thread-local last error (api.hip)
}

namespace holoflow::core {
namespace {

using curaii::HfxError;

[[noreturn]] void fail(const char *who, const std::string &msg) {
  throw std::invalid_argument(std::string(who) + " inference error: " + msg);
}
void check(bool c, const char *who, const std::string &msg) {
  if (!c) fail(who, msg);
}

int norm_axis(long long a, size_t rank, const char *who) {
  const long long r = a < 0 ? a + (long long)rank : a;
  check(r >= 0 && r < (long long)rank, who, "axis out of bounds");
  return (int)r;
}

size_t prod(const std::vector<size_t> &s, size_t b, size_t e) {
  size_t p = 1;
  for (size_t i = b; i < e && i < s.size(); ++i) p *= s[i];
  return p;
}

hfx_stream_t hs(hipStream_t s) { return reinterpret_cast<hfx_stream_t>(s); }

// Reference settings store lambda/dx/dy/z as float: round through float first.
double f32(const Json &j, const char *key) { return (double)(float)j.at(key).number(); }

InferResult sync_result(std::vector<TDesc> in, std::vector<TDesc> out, std::vector<InPlace> ip = {}) {
  InferResult r;
  r.owned_inputs.assign(in.size(), false);
  r.owned_outputs.assign(out.size(), false);
  r.input_descs = std::move(in);
  r.output_descs = std::move(out);
  r.in_place = std::move(ip);
  r.kind = TaskKind::Sync;
  return r;
}

void require_device(const TDesc &d, const char *who) {
  check(d.mem_loc == MemLoc::Device, who, "input must be in device memory");
}

// ================================================================== Conversion
int hfx_code(DType d) {
  switch (d) {
    case DType::U8: return HFX_DTYPE_U8;
    case DType::U16: return HFX_DTYPE_U16;
    case DType::F32: return HFX_DTYPE_F32;
    case DType::CF32: return HFX_DTYPE_CF32;
  }
  return 0;
}

// (input dtype, target, strategy) -> output dtype: the reference's table (conversion.cu:
// 138-150) plus U16 -> F32 Real, which the path's graph uses (graph_builder.cc:242).
bool conversion_out(DType in, DType target, int strategy, DType &out) {
  struct Row {
    DType in, target;
    int strategy;
  };
  static const Row rows[] = {
      {DType::U8, DType::F32, HFX_CONV_REAL},      {DType::U8, DType::CF32, HFX_CONV_REAL},
      {DType::U16, DType::F32, HFX_CONV_REAL},     {DType::U16, DType::CF32, HFX_CONV_REAL},
      {DType::F32, DType::U8, HFX_CONV_SCALED},    {DType::F32, DType::U16, HFX_CONV_SCALED},
      {DType::F32, DType::CF32, HFX_CONV_REAL},    {DType::CF32, DType::F32, HFX_CONV_MODULUS},
      {DType::CF32, DType::F32, HFX_CONV_ARGUMENT},
  };
  for (const auto &r : rows)
    if (r.in == in && r.target == target && r.strategy == strategy) {
      out = target;
      return true;
    }
  return false;
}

int strategy_code(const std::string &s) {
  if (s == "Real") return HFX_CONV_REAL;
  if (s == "Scaled") return HFX_CONV_SCALED;
  if (s == "Modulus") return HFX_CONV_MODULUS;
  if (s == "Argument") return HFX_CONV_ARGUMENT;
  throw std::invalid_argument("Unknown strategy type for conversion: " + s);
}

class ConversionTask : public ISyncTask {
 public:
  hipStream_t st;
  size_t n;
  int in_dt, out_dt, strategy;
  curaii::unique_device_ptr<std::byte> scratch;
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_convert(c.inputs[0].data(), in_dt, c.outputs[0].data(), out_dt, strategy, (int64_t)n,
                          scratch.get(), hs(st)));
    return OpResult::Ok;
  }
};

class ConversionFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = "ConversionFactory";
    const std::string ts = s.at("target").str();
    if (ts != "U8" && ts != "U16" && ts != "F32" && ts != "CF32")
      throw std::invalid_argument("Unknown target type for conversion: " + ts);
    const DType target = dtype_from_string(ts);
    const int strategy = strategy_code(s.at("strategy").str());
    check(in.size() == 1, who, "Expected exactly one input tensor");
    check(in[0].mem_loc == MemLoc::Device, who, "Input tensor must be in device memory");
    check(in[0].is_contiguous(), who, "Input tensor must be C-contiguous");
    DType out;
    check(conversion_out(in[0].dtype, target, strategy, out), who, "Unsupported conversion configuration");
    return sync_result({in[0]}, {TDesc(in[0].shape, out, MemLoc::Device)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    const auto r = infer(in, s);
    auto t = std::make_unique<ConversionTask>();
    t->st = ctx.stream;
    t->n = in[0].num_elements();
    t->in_dt = hfx_code(in[0].dtype);
    t->out_dt = hfx_code(r.output_descs[0].dtype);
    t->strategy = strategy_code(s.at("strategy").str());
    if (t->strategy == HFX_CONV_SCALED)
      t->scratch = curaii::make_unique_device_ptr<std::byte>((size_t)hfx_convert_scratch_bytes((int64_t)t->n));
    return t;
  }
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &s,
                                    const SyncCreateCtx &ctx) const override {
    (void)infer(in, s);
    auto *t = dynamic_cast<ConversionTask *>(old.get());
    if (t && t->n == in[0].num_elements() && t->in_dt == hfx_code(in[0].dtype) &&
        t->strategy == strategy_code(s.at("strategy").str()) &&
        t->out_dt == hfx_code(dtype_from_string(s.at("target").str()))) {
      t->st = ctx.stream;
      return old;
    }
    return create(in, s, ctx);
  }
};

// ================================================================== RFFT / FFT
double fft_scale(const Json &s, int n, const char *who) {
  const std::string norm = s.contains("norm") ? s.at("norm").str() : "backward";
  if (norm == "backward") return 1.0;
  if (norm == "forward") return 1.0 / n;
  if (norm == "ortho") return 1.0 / std::sqrt((double)n);
  fail(who, "unknown norm '" + norm + "'");
}

class DftTask : public ISyncTask {
 public:
  hipStream_t st;
  int dtype, n, kb, ke;
  long long outer, inner;
  float scale;
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_dft_axis(c.inputs[0].data(), dtype, outer, n, inner, kb, ke, scale, c.outputs[0].data(), hs(st)));
    return OpResult::Ok;
  }
};

template <bool REAL> class DftFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = REAL ? "RFFTFactory" : "FFTFactory";
    check(in.size() == 1, who, "expected exactly one input");
    require_device(in[0], who);
    check(in[0].is_contiguous(), who, "input must be contiguous");
    check(in[0].rank() >= 1, who, "input must have rank >= 1");
    const DType d = in[0].dtype;
    check(d == DType::F32 || d == DType::U16 || (!REAL && d == DType::CF32), who,
          REAL ? "input must be F32 (or U16)" : "input must be F32, CF32 (or U16)");
    const int ax = norm_axis(s.get_int("axis", -1), in[0].rank(), who);
    (void)fft_scale(s, (int)in[0].shape[ax], who);
    auto shape = in[0].shape;
    if (REAL) shape[ax] = shape[ax] / 2 + 1;
    return sync_result({in[0]}, {TDesc(shape, DType::CF32, in[0].mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    const auto r = infer(in, s);
    const int ax = norm_axis(s.get_int("axis", -1), in[0].rank(), "DftFactory");
    auto t = std::make_unique<DftTask>();
    t->st = ctx.stream;
    t->dtype = in[0].dtype == DType::U16 ? 1 : in[0].dtype == DType::F32 ? 2 : 3;
    t->n = (int)in[0].shape[ax];
    t->kb = 0;
    t->ke = (int)r.output_descs[0].shape[ax];
    t->outer = (long long)prod(in[0].shape, 0, ax);
    t->inner = (long long)prod(in[0].shape, ax + 1, in[0].rank());
    t->scale = (float)fft_scale(s, t->n, "DftFactory");
    return t;
  }
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &s,
                                    const SyncCreateCtx &ctx) const override {
    auto *t = dynamic_cast<DftTask *>(old.get());
    if (t) {
      auto fresh = create(in, s, ctx);
      auto *f = static_cast<DftTask *>(fresh.get());
      if (f->dtype == t->dtype && f->n == t->n && f->ke == t->ke && f->outer == t->outer &&
          f->inner == t->inner && f->scale == t->scale) {
        t->st = ctx.stream;
        return old;
      }
      return fresh;
    }
    return create(in, s, ctx);
  }
};

// ================================================================== Copy / Concatenate
void copy_view(const TView &src, void *dst, hipStream_t st) {
  const TDesc &d = src.desc;
  std::vector<int64_t> shape(d.shape.begin(), d.shape.end()), strides(d.strides.begin(), d.strides.end());
  if (shape.empty()) {
    shape.push_back(1);
    strides.push_back((int64_t)size_of(d.dtype));
  }
  HFX_CHECK(hfx_strided_copy(src.storage->ptr, shape.data(), strides.data(), (int64_t)d.offset,
                             (int32_t)shape.size(), (int32_t)size_of(d.dtype), dst, 0, 0, hs(st)));
}

class CopyTask : public ISyncTask {
 public:
  hipStream_t st;
  OpResult execute(SyncCtx &c) override {
    copy_view(c.inputs[0], c.outputs[0].data(), st);
    return OpResult::Ok;
  }
};

class CopyFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &) const override {
    check(in.size() == 1, "CopyFactory", "expected exactly one input");
    require_device(in[0], "CopyFactory");
    return sync_result({in[0]}, {TDesc(in[0].shape, in[0].dtype, in[0].mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    (void)infer(in, s);
    auto t = std::make_unique<CopyTask>();
    t->st = ctx.stream;
    return t;
  }
};

// axis < 0: inputs flattened (settings "axis": null); otherwise the normalised axis.
int concat_axis(const Json &s, size_t rank, const char *who) {
  if (!s.contains("axis")) return norm_axis(0, rank, who);
  const Json &a = s.at("axis");
  if (a.is_null()) return -1;
  if (!a.is_number()) throw std::invalid_argument("ConcatenateSettings: axis must be int or null");
  return norm_axis(a.integer(), rank, who);
}

// Each input is one strided gather into its band of the output: [outer][n_i*inner] rows
// at pitch N*inner (N = output extent along the axis), offset by the elements before it.
class ConcatenateTask : public ISyncTask {
 public:
  hipStream_t st;
  int axis;
  std::vector<TDesc> in;
  TDesc out;
  OpResult execute(SyncCtx &c) override {
    const size_t esz = size_of(out.dtype);
    const size_t inner = axis < 0 ? 1 : prod(out.shape, axis + 1, out.rank());
    const size_t pitch = axis < 0 ? 0 : out.shape[axis] * inner;
    size_t off = 0;
    for (size_t i = 0; i < in.size(); ++i) {
      const TDesc &d = in[i];
      const size_t n_i = axis < 0 ? d.num_elements() : d.shape[axis];
      std::vector<int64_t> shape(d.shape.begin(), d.shape.end()), strides(d.strides.begin(), d.strides.end());
      if (shape.empty()) {
        shape.push_back(1);
        strides.push_back((int64_t)esz);
      }
      std::byte *dst = c.outputs[0].data() + off * inner * esz;
      const int64_t block = axis < 0 ? 0 : (int64_t)(n_i * inner);
      HFX_CHECK(hfx_strided_copy(c.inputs[i].storage->ptr, shape.data(), strides.data(), (int64_t)d.offset,
                                 (int32_t)shape.size(), (int32_t)esz, dst, block, (int64_t)pitch, hs(st)));
      off += n_i;
    }
    return OpResult::Ok;
  }
};

class ConcatenateFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = "Concatenate";
    check(!in.empty(), who, "expected at least 1 input");
    check(in[0].mem_loc == MemLoc::Device, who, "only Device tensors are supported");
    for (const auto &d : in) {
      check(d.mem_loc == in[0].mem_loc, who, "all inputs must share the same mem_loc");
      check(d.dtype == in[0].dtype, who, "all inputs must share the same dtype");
      check(d.rank() <= 8, who, "input ndim too large");
    }
    const size_t rank = in[0].rank();
    if (s.contains("axis") && s.at("axis").is_null()) {
      size_t total = 0;
      for (const auto &d : in) total += d.num_elements();
      return sync_result(std::vector<TDesc>(in.begin(), in.end()), {TDesc({total}, in[0].dtype, in[0].mem_loc)});
    }
    check(rank > 0, who, "input ndim must be > 0");
    for (const auto &d : in) check(d.rank() == rank, who, "all inputs must share the same rank");
    const int ax = concat_axis(s, rank, who);
    auto shape = in[0].shape;
    shape[ax] = 0;
    for (const auto &d : in) {
      for (size_t k = 0; k < rank; ++k)
        if ((int)k != ax) check(d.shape[k] == in[0].shape[k], who, "all input shapes must match except along axis");
      shape[ax] += d.shape[ax];
    }
    return sync_result(std::vector<TDesc>(in.begin(), in.end()), {TDesc(shape, in[0].dtype, in[0].mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    const auto r = infer(in, s);
    auto t = std::make_unique<ConcatenateTask>();
    t->st = ctx.stream;
    t->axis = (s.contains("axis") && s.at("axis").is_null()) ? -1 : concat_axis(s, in[0].rank(), "Concatenate");
    t->in.assign(in.begin(), in.end());
    t->out = r.output_descs[0];
    return t;
  }
};

// ================================================================== views
class NoopTask : public ISyncTask {
 public:
  OpResult execute(SyncCtx &) override { return OpResult::Ok; }
};

std::vector<size_t> resolve_shape(const Json &req_json, size_t total) {
  std::vector<long long> req;
  for (const auto &v : req_json.items()) req.push_back(v.integer());
  long long known = 1;
  int infer_at = -1;
  for (size_t i = 0; i < req.size(); ++i) {
    if (req[i] == -1) {
      if (infer_at >= 0) throw std::invalid_argument("Reshape: Only one dimension can be -1");
      infer_at = (int)i;
    } else {
      if (req[i] < 0) throw std::invalid_argument("Reshape: Dimensions cannot be negative (except -1)");
      known *= req[i];
    }
  }
  if (infer_at >= 0) {
    if (known == 0 || (long long)total % known != 0)
      throw std::invalid_argument("Reshape: Total elements must be divisible by known dimensions");
    req[infer_at] = (long long)total / known;
  }
  std::vector<size_t> shape(req.begin(), req.end());
  if (prod(shape, 0, shape.size()) != total)
    throw std::invalid_argument("Reshape: Total elements mismatch between input and target shape");
  return shape;
}

// Byte strides that present `in` as `out_shape` without moving data, or nothing when the
// input's layout does not allow it.  Input dims (size-1 dims dropped) are merged into
// runs that are contiguous with each other; every run must be covered by a whole number
// of output dims, whose strides are then laid out inside that run.
std::optional<std::vector<size_t>> view_strides(const TDesc &in, const std::vector<size_t> &out_shape) {
  const size_t esz = size_of(in.dtype);
  std::vector<size_t> shp, str;
  for (size_t i = 0; i < in.rank(); ++i)
    if (in.shape[i] != 1) {
      shp.push_back(in.shape[i]);
      str.push_back(in.strides[i]);
    }
  std::vector<size_t> res(out_shape.size(), esz);
  if (prod(out_shape, 0, out_shape.size()) == 0) return TDesc::contiguous_strides(out_shape, in.dtype);
  size_t oi = 0;
  size_t ii = 0;
  while (ii < shp.size()) {
    // run of input dims [ii, ie) contiguous with each other
    size_t ie = ii + 1;
    while (ie < shp.size() && str[ie - 1] == str[ie] * shp[ie]) ++ie;
    const size_t run = prod(shp, ii, ie), run_stride = str[ie - 1];
    // output dims covering this run exactly
    size_t acc = 1, ob = oi;
    while (oi < out_shape.size() && acc < run) acc *= out_shape[oi++];
    while (oi < out_shape.size() && out_shape[oi] == 1) ++oi;  // trailing unit dims join the run
    if (acc != run) return std::nullopt;
    size_t st = run_stride;
    for (size_t k = oi; k-- > ob;) {
      res[k] = st;
      st *= out_shape[k];
    }
    ii = ie;
  }
  for (; oi < out_shape.size(); ++oi) res[oi] = esz;  // only unit dims remain
  return res;
}

class ReshapeCopyTask : public ISyncTask {
 public:
  hipStream_t st;
  OpResult execute(SyncCtx &c) override;
};

class ReshapeFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    if (in.size() != 1) throw std::runtime_error("Reshape requires exactly 1 input");
    const TDesc &src = in[0];
    const auto shape = resolve_shape(s.at("shape"), src.num_elements());
    const auto vs = view_strides(src, shape);
    bool copy = !vs.has_value();
    if (s.contains("copy") && !s.at("copy").is_null()) {
      copy = s.at("copy").boolean();
      if (!copy && !vs) throw std::invalid_argument("Reshape: copy=false requested, but tensor layout prevents a view.");
    }
    if (copy) return sync_result({src}, {TDesc(shape, src.dtype, src.mem_loc)});
    TDesc out = src;
    out.shape = shape;
    out.strides = *vs;
    return sync_result({src}, {out}, {{0, 0}});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    const auto r = infer(in, s);
    if (!r.in_place.empty()) return std::make_unique<NoopTask>();
    require_device(in[0], "ReshapeFactory");
    auto t = std::make_unique<ReshapeCopyTask>();
    t->st = ctx.stream;
    return t;
  }
};

OpResult ReshapeCopyTask::execute(SyncCtx &c) {
  copy_view(c.inputs[0], c.outputs[0].data(), st);
  return OpResult::Ok;
}

class SliceFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = "Slice";
    check(in.size() == 1, who, "expected exactly 1 input");
    const TDesc &d = in[0];
    const int ndim = (int)d.rank();
    check(ndim > 0, who, "input ndim must be > 0");
    const auto &items = s.at("slices").items();
    check((int)items.size() == ndim, who, "number of slice items must match input ndim");
    std::vector<size_t> shape, strides;
    size_t offset = d.offset;
    for (int i = 0; i < ndim; ++i) {
      const Json &it = items[i];
      const long long dim = (long long)d.shape[i];
      if (it.is_number()) {  // integer index: drops the axis
        check(dim > 0, who, "cannot index into 0-sized dimension");
        long long idx = it.integer();
        if (idx < 0) idx += dim;
        if (idx < 0 || idx >= dim)
          throw std::out_of_range("Slice index " + std::to_string(idx) + " is out of bounds for dimension size " +
                                  std::to_string(dim));
        offset += (size_t)idx * d.strides[i];
        continue;
      }
      const bool has_start = it.contains("start") && !it.at("start").is_null();
      const bool has_stop = it.contains("stop") && !it.at("stop").is_null();
      long long start = has_start ? it.at("start").integer() : 0;
      long long stop = has_stop ? it.at("stop").integer() : dim;
      const long long step = it.get_int("step", 1);
      check(step > 0, who, "only positive step is supported for now");
      if (start < 0) start += dim;
      if (stop < 0) stop += dim;
      start = std::clamp(start, 0LL, dim);
      stop = std::clamp(stop, 0LL, dim);
      const long long len = start >= stop ? 0 : (stop - start + step - 1) / step;
      offset += (size_t)start * d.strides[i];
      shape.push_back((size_t)len);
      strides.push_back(d.strides[i] * (size_t)step);
    }
    return sync_result({d}, {TDesc(shape, d.dtype, d.mem_loc, strides, offset)}, {{0, 0}});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &) const override {
    (void)infer(in, s);
    return std::make_unique<NoopTask>();
  }
};

// ================================================================== FresnelDiffraction
struct FresnelSettings {
  double lambda, dx, dy, z;
  bool skip_phase_shift, output_magnitude;
  std::vector<long long> axes;
  bool operator==(const FresnelSettings &o) const {
    return lambda == o.lambda && dx == o.dx && dy == o.dy && z == o.z && skip_phase_shift == o.skip_phase_shift &&
           output_magnitude == o.output_magnitude && axes == o.axes;
  }
};

FresnelSettings fresnel_settings(const Json &j) {
  FresnelSettings s{f32(j, "lambda"), f32(j, "dx"), f32(j, "dy"), f32(j, "z"), j.get_bool("skip_phase_shift", false),
                    j.get_bool("output_magnitude", false), {-2, -1}};
  if (j.contains("axes") && !j.at("axes").is_null()) {
    s.axes.clear();
    for (const auto &a : j.at("axes").items()) s.axes.push_back(a.integer());
  }
  return s;
}

class FresnelTask : public ISyncTask {
 public:
  hipStream_t st;
  FresnelSettings s;
  TDesc idesc;
  hfx_fresnel *plan = nullptr;
  ~FresnelTask() override {
    if (plan) hfx_fresnel_destroy(plan);
  }
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_fresnel_execute(plan, c.inputs[0].data(), c.outputs[0].data(), hs(st)));
    return OpResult::Ok;
  }
};

class FresnelDiffractionFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &j) const override {
    const char *who = "FresnelDiffractionFactory";
    const auto s = fresnel_settings(j);
    check(in.size() == 1, who, "expected exactly one input");
    const TDesc &d = in[0];
    check(d.rank() >= 2, who, "input must be a tensor of rank 2 or higher");
    check(d.dtype == DType::CF32 || d.dtype == DType::F32, who, "input must be complex float32 or real float32");
    require_device(d, who);
    check(s.axes.size() == 2, who, "axes must contain exactly two dimensions");
    const int a0 = norm_axis(s.axes[0], d.rank(), who), a1 = norm_axis(s.axes[1], d.rank(), who);
    check(a0 != a1, who, "axes must be distinct");
    check(a0 == (int)d.rank() - 2 && a1 == (int)d.rank() - 1, who, "only the two trailing axes are supported");
    check(d.is_contiguous(), who, "input must be contiguous");
    check(s.lambda > 0, who, "wavelength must be positive");
    check(s.dx > 0, who, "dx must be positive");
    check(s.dy > 0, who, "dy must be positive");
    check(s.dx == s.dy, who, "dx must equal dy");
    const DType od = s.output_magnitude ? DType::F32 : DType::CF32;
    return sync_result({d}, {TDesc(d.shape, od, d.mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &j, const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto t = std::make_unique<FresnelTask>();
    t->st = ctx.stream;
    t->s = fresnel_settings(j);
    t->idesc = in[0];
    const size_t r = in[0].rank();
    HFX_CHECK(hfx_fresnel_create((int64_t)prod(in[0].shape, 0, r - 2), (int32_t)in[0].shape[r - 2],
                                 (int32_t)in[0].shape[r - 1], t->s.lambda, t->s.dx, t->s.dy, t->s.z,
                                 in[0].dtype == DType::F32, t->s.output_magnitude, t->s.skip_phase_shift,
                                 hs(ctx.stream), &t->plan));
    return t;
  }
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &j,
                                    const SyncCreateCtx &ctx) const override {
    auto *t = dynamic_cast<FresnelTask *>(old.get());
    if (!t || in.size() != 1) return create(in, j, ctx);
    const auto s = fresnel_settings(j);
    if (in[0] != t->idesc) return create(in, j, ctx);
    auto zless = s;
    zless.z = t->s.z;
    if (s == t->s) {
      t->st = ctx.stream;
      return old;
    }
    if (zless == t->s) {  // z-only change: rewrite the lens in place
      (void)infer(in, j);
      t->st = ctx.stream;
      HFX_CHECK(hfx_fresnel_update_z(t->plan, s.z, hs(ctx.stream)));
      t->s = s;
      return old;
    }
    return create(in, j, ctx);
  }
};

// ================================================================== AngularSpectrum
hfx_asm_desc asm_desc(const Json &j, const TDesc &d) {
  hfx_asm_desc a{};
  const size_t r = d.rank();
  a.batch = (int64_t)prod(d.shape, 0, r - 2);
  a.height = (int32_t)d.shape[r - 2];
  a.width = (int32_t)d.shape[r - 1];
  a.lambda_m = f32(j, "lambda");
  a.dx_m = f32(j, "dx");
  a.dy_m = f32(j, "dy");
  a.z_m = f32(j, "z");
  a.in_real = d.dtype == DType::F32;
  if (j.contains("filter") && !j.at("filter").is_null()) {
    const auto &f = j.at("filter");
    a.filter_enabled = 1;
    a.r_inner = (double)(uint32_t)f.at("r_inner").integer();
    a.r_outer = (double)(uint32_t)f.at("r_outer").integer();
    a.s_inner = (double)(uint32_t)f.at("s_inner").integer();
    a.s_outer = (double)(uint32_t)f.at("s_outer").integer();
  }
  if (j.contains("padding") && !j.at("padding").is_null()) {
    a.pad_width = (int32_t)j.at("padding").at("width").integer();
    a.pad_height = (int32_t)j.at("padding").at("height").integer();
  }
  return a;
}

class AsmTask : public ISyncTask {
 public:
  hipStream_t st;
  Json settings;
  TDesc idesc;
  hfx_asm *plan = nullptr;
  ~AsmTask() override {
    if (plan) hfx_asm_destroy(plan);
  }
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_asm_execute(plan, c.inputs[0].data(), c.outputs[0].data(), hs(st)));
    return OpResult::Ok;
  }
};

class AngularSpectrumFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &j) const override {
    const char *who = "AngularSpectrumFactory";
    check(in.size() == 1, who, "expected exactly one input");
    const TDesc &d = in[0];
    check(d.rank() >= 2, who, "input must be a tensor of rank 2 or higher");
    check(d.dtype == DType::CF32 || d.dtype == DType::F32, who, "input must be complex float32 or real float32");
    require_device(d, who);
    check(d.is_contiguous(), who, "input must be C-contiguous");
    const auto a = asm_desc(j, d);
    check(a.lambda_m > 0, who, "wavelength must be positive");
    check(a.dx_m > 0, who, "dx must be positive");
    check(a.dy_m > 0, who, "dy must be positive");
    check(a.dx_m == a.dy_m, who, "dx must equal dy");
    auto shape = d.shape;
    const bool padded = j.contains("padding") && !j.at("padding").is_null();
    if (padded) {
      check(a.pad_width > 0, who, "padding width must be positive");
      check(a.pad_height > 0, who, "padding height must be positive");
      check(a.pad_width >= a.width, who, "padding width must be at least the input width");
      check(a.pad_height >= a.height, who, "padding height must be at least the input height");
      check((a.pad_width - a.width) % 2 == 0, who, "horizontal padding must be even");
      check((a.pad_height - a.height) % 2 == 0, who, "vertical padding must be even");
      shape[shape.size() - 2] = (size_t)a.pad_height;
      shape[shape.size() - 1] = (size_t)a.pad_width;
    }
    std::vector<InPlace> ip;
    if (d.dtype == DType::CF32 && !padded) ip.push_back({0, 0});
    return sync_result({d}, {TDesc(shape, DType::CF32, d.mem_loc)}, ip);
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &j, const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto t = std::make_unique<AsmTask>();
    t->st = ctx.stream;
    t->settings = j;
    t->idesc = in[0];
    const auto a = asm_desc(j, in[0]);
    HFX_CHECK(hfx_asm_create(&a, hs(ctx.stream), &t->plan));
    return t;
  }
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &j,
                                    const SyncCreateCtx &ctx) const override {
    auto *t = dynamic_cast<AsmTask *>(old.get());
    if (t && in.size() == 1 && in[0] == t->idesc && j == t->settings) {
      t->st = ctx.stream;
      return old;
    }
    return create(in, j, ctx);
  }
};

// ================================================================== Filter2D / Flatfield
// Optional in-path stages (filter2d.cu:359-497, flatfield.cu:323-451): settings keys,
// checks and reuse rules as the reference's factories; the arithmetic runs through
// hfx_filter2d_* / hfx_flatfield_* (2-D FFT passes, or the spatial blur).
struct StageGeo {
  int64_t batch;
  int32_t H, W;
};

StageGeo stage_geo(const TDesc &d) {
  const size_t r = d.rank();
  return {(int64_t)prod(d.shape, 0, r - 2), (int32_t)d.shape[r - 2], (int32_t)d.shape[r - 1]};
}

class Filter2DTask : public ISyncTask {
 public:
  hipStream_t st;
  Json settings;
  TDesc idesc;
  hfx_filter2d *plan = nullptr;
  ~Filter2DTask() override {
    if (plan) hfx_filter2d_destroy(plan);
  }
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_filter2d_execute(plan, c.inputs[0].data(), c.outputs[0].data(), hs(st)));
    return OpResult::Ok;
  }
};

class Filter2DFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &j) const override {
    const char *who = "Filter2DFactory";
    const long long ri = j.at("r_inner").integer(), ro = j.at("r_outer").integer(), si = j.at("s_inner").integer(),
                    so = j.at("s_outer").integer();
    check(in.size() == 1, who, "expected exactly one input");
    const TDesc &d = in[0];
    check(d.rank() >= 2, who, "input must be a tensor of rank 2 or higher");
    check(d.dtype == DType::CF32 || d.dtype == DType::F32, who, "input must be complex float32 or real float32");
    require_device(d, who);
    check(d.is_contiguous(), who, "input must be C-contiguous");
    check(ri >= 0, who, "r_inner must be non-negative");
    check(ro >= 0, who, "r_outer must be non-negative");
    check(si >= 0, who, "s_inner must be non-negative");
    check(so >= 0, who, "s_outer must be non-negative");
    std::vector<InPlace> ip;
    if (d.dtype == DType::CF32) ip.push_back({0, 0});
    return sync_result({d}, {d}, ip);
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &j, const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto t = std::make_unique<Filter2DTask>();
    t->st = ctx.stream;
    t->settings = j;
    t->idesc = in[0];
    const auto g = stage_geo(in[0]);
    HFX_CHECK(hfx_filter2d_create(g.batch, g.H, g.W, in[0].dtype == DType::F32, (int32_t)j.at("r_inner").integer(),
                                  (int32_t)j.at("r_outer").integer(), (int32_t)j.at("s_inner").integer(),
                                  (int32_t)j.at("s_outer").integer(), hs(ctx.stream), &t->plan));
    return t;
  }
  // Same settings and input layout keep the task (filter2d.cu:471-497).
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &j,
                                    const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto *t = dynamic_cast<Filter2DTask *>(old.get());
    if (t && in[0] == t->idesc && j == t->settings) {
      t->st = ctx.stream;
      return old;
    }
    return create(in, j, ctx);
  }
};

class FlatfieldTask : public ISyncTask {
 public:
  hipStream_t st;
  Json settings;
  TDesc idesc;
  hfx_flatfield *plan = nullptr;
  ~FlatfieldTask() override {
    if (plan) hfx_flatfield_destroy(plan);
  }
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_flatfield_execute(plan, reinterpret_cast<const float *>(c.inputs[0].data()),
                                    reinterpret_cast<float *>(c.outputs[0].data()), hs(st)));
    return OpResult::Ok;
  }
};

class FlatfieldFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &j) const override {
    const char *who = "FlatfieldFactory";
    const float sy = (float)j.at("sigma_y").number(), sx = (float)j.at("sigma_x").number();
    check(in.size() == 1, who, "expected exactly one input");
    const TDesc &d = in[0];
    check(d.rank() >= 2, who, "input must be a tensor of rank 2 or higher");
    check(d.dtype == DType::F32, who, "input must be float32");
    require_device(d, who);
    check(d.is_contiguous(), who, "input must be C-contiguous");
    check(d.num_elements() > 0, who, "input must not be empty");
    check(sy > 0.0f, who, "sigma_y must be positive");
    check(sx > 0.0f, who, "sigma_x must be positive");
    return sync_result({d}, {d});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &j, const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto t = std::make_unique<FlatfieldTask>();
    t->st = ctx.stream;
    t->settings = j;
    t->idesc = in[0];
    const auto g = stage_geo(in[0]);
    const hfx_status rc = hfx_flatfield_create(g.batch, g.H, g.W, (float)j.at("sigma_y").number(),
                                               (float)j.at("sigma_x").number(), hs(ctx.stream), &t->plan, nullptr);
    if (rc == HFX_E_INVALID) fail("FlatfieldFactory", hfx_last_error());
    HFX_CHECK(rc);
    return t;
  }
  // Same settings and input layout keep the task (flatfield.cu:425-451).
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &j,
                                    const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto *t = dynamic_cast<FlatfieldTask *>(old.get());
    if (t && in[0] == t->idesc && j == t->settings) {
      t->st = ctx.stream;
      return old;
    }
    return create(in, j, ctx);
  }
};

// ================================================================== reductions
struct ReduceGeo {
  size_t outer, planes, inner;
};

ReduceGeo reduce_geo(const TDesc &d, const Json &s, const char *who, TDesc &out) {
  std::vector<long long> axes;
  if (s.contains("axis") && !s.at("axis").is_null()) {
    const auto &a = s.at("axis");
    if (a.is_array())
      for (const auto &v : a.items()) axes.push_back(v.integer());
    else
      axes.push_back(a.integer());
  } else if (s.contains("axes") && !s.at("axes").is_null()) {
    for (const auto &v : s.at("axes").items()) axes.push_back(v.integer());
  }
  const bool keep = s.get_bool("keepdims", false);
  std::vector<bool> red(d.rank(), axes.empty());
  for (long long a : axes) red[norm_axis(a, d.rank(), who)] = true;
  // the reduced axes must form one contiguous run (the reductions on the path reduce one axis)
  int first = -1, last = -1;
  for (int i = 0; i < (int)d.rank(); ++i)
    if (red[i]) {
      if (first < 0) first = i;
      last = i;
    }
  check(first >= 0, who, "no axis to reduce");
  for (int i = first; i <= last; ++i) check(red[i], who, "reduced axes must be adjacent");
  std::vector<size_t> shape;
  for (size_t i = 0; i < d.rank(); ++i) {
    if (!red[i])
      shape.push_back(d.shape[i]);
    else if (keep)
      shape.push_back(1);
  }
  out = TDesc(shape, DType::F32, d.mem_loc);
  ReduceGeo g{prod(d.shape, 0, first), prod(d.shape, first, last + 1), prod(d.shape, last + 1, d.rank())};
  check(g.planes > 0 && g.outer * g.inner > 0, who, "reduction has zero elements");
  return g;
}

class ReduceTask : public ISyncTask {
 public:
  hipStream_t st;
  ReduceGeo g;
  int complex_in;
  uint32_t flags;
  size_t esz;
  OpResult execute(SyncCtx &c) override {
    for (size_t o = 0; o < g.outer; ++o)
      HFX_CHECK(hfx_weighted_mean(c.inputs[0].data() + o * g.planes * g.inner * esz, complex_in, (int64_t)g.planes,
                                  (int64_t)g.inner, nullptr, flags,
                                  reinterpret_cast<float *>(c.outputs[0].data()) + o * g.inner, hs(st)));
    return OpResult::Ok;
  }
};

// ABS: MeanAbs (|x| of real input too); otherwise Mean (real input only).
template <bool ABS> class ReduceFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = ABS ? "MeanAbsFactory" : "MeanFactory";
    check(in.size() == 1, who, "expected exactly one input");
    require_device(in[0], who);
    check(in[0].is_contiguous(), who, "input must be contiguous");
    if (ABS)
      check(in[0].dtype == DType::CF32 || in[0].dtype == DType::F32, who, "input must be CF32 or F32");
    else
      check(in[0].dtype == DType::F32, who, "only F32 input is supported");
    TDesc out;
    (void)reduce_geo(in[0], s, who, out);
    return sync_result({in[0]}, {out});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    (void)infer(in, s);
    auto t = std::make_unique<ReduceTask>();
    TDesc out;
    t->g = reduce_geo(in[0], s, "ReduceFactory", out);
    t->st = ctx.stream;
    t->complex_in = in[0].dtype == DType::CF32;
    t->flags = ABS ? HFX_REDUCE_ABS : 0u;
    t->esz = size_of(in[0].dtype);
    return t;
  }
};

class AbsFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &) const override {
    const char *who = "AbsFactory";
    check(in.size() == 1, who, "expected exactly one input");
    require_device(in[0], who);
    check(in[0].is_contiguous(), who, "input must be contiguous");
    check(in[0].dtype == DType::CF32 || in[0].dtype == DType::F32, who, "input must be CF32 or F32");
    return sync_result({in[0]}, {TDesc(in[0].shape, DType::F32, in[0].mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    (void)infer(in, s);
    auto t = std::make_unique<ReduceTask>();  // |x| = mean of one plane
    t->g = ReduceGeo{1, 1, in[0].num_elements()};
    t->st = ctx.stream;
    t->complex_in = in[0].dtype == DType::CF32;
    t->flags = HFX_REDUCE_ABS;
    t->esz = size_of(in[0].dtype);
    return t;
  }
};

// ================================================================== Arange / Multiply
class ArangeTask : public ISyncTask {
 public:
  hipStream_t st;
  curaii::unique_host_ptr<std::byte> host;
  size_t bytes;
  OpResult execute(SyncCtx &c) override {
    if (bytes) HIP_CHECK(hipMemcpyAsync(c.outputs[0].data(), host.get(), bytes, hipMemcpyHostToDevice, st));
    return OpResult::Ok;
  }
};

class ArangeFactory : public ISyncTaskFactory {
 public:
  static size_t length(double start, double stop, double step) {
    const bool fwd = step > 0.0;
    if ((fwd && stop <= start) || (!fwd && stop >= start)) return 0;
    const double n = std::ceil((stop - start) / step);
    if (!std::isfinite(n) || n <= 0.0) throw std::invalid_argument("Arange: invalid resulting length");
    if (n > 1.8e19) throw std::invalid_argument("Arange: resulting length too large");
    return (size_t)n;
  }
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = "ArangeFactory";
    const DType dt = s.contains("dtype") ? dtype_from_string(s.at("dtype").str()) : DType::F32;
    const bool host = s.contains("device") && s.at("device").str() == "Host";
    check(in.empty(), who, "expected zero inputs");
    check(!host, who, "only Device output is supported (for now)");
    const double step = s.at("step").number();
    check(step != 0.0, who, "step must be non-zero");
    const size_t n = length(s.at("start").number(), s.at("stop").number(), step);
    return sync_result({}, {TDesc({n}, dt, MemLoc::Device)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    const auto r = infer(in, s);
    const size_t n = r.output_descs[0].shape[0];
    const DType dt = r.output_descs[0].dtype;
    auto t = std::make_unique<ArangeTask>();
    t->st = ctx.stream;
    t->bytes = n * size_of(dt);
    t->host = curaii::make_unique_host_ptr<std::byte>(t->bytes);
    const double start = s.at("start").number(), step = s.at("step").number();
    std::byte *h = t->host.get();
    for (size_t i = 0; i < n; ++i) {  // start + step*i in double, then cast (arange.cu:61-76)
      const double v = start + step * (double)i;
      switch (dt) {
        case DType::U8: reinterpret_cast<uint8_t *>(h)[i] = (uint8_t)v; break;
        case DType::U16: reinterpret_cast<uint16_t *>(h)[i] = (uint16_t)v; break;
        case DType::F32: reinterpret_cast<float *>(h)[i] = (float)v; break;
        case DType::CF32:
          reinterpret_cast<float *>(h)[2 * i] = (float)v;
          reinterpret_cast<float *>(h)[2 * i + 1] = 0.f;
          break;
      }
    }
    return t;
  }
};

class MultiplyTask : public ISyncTask {
 public:
  hipStream_t st;
  std::vector<int64_t> shape, sa, sb;
  int ac, bc;
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_multiply(c.inputs[0].data(), ac, sa.data(), c.inputs[1].data(), bc, sb.data(), (int32_t)shape.size(),
                           shape.data(), c.outputs[0].data(), hs(st)));
    return OpResult::Ok;
  }
};

class MultiplyFactory : public ISyncTaskFactory {
 public:
  // NumPy broadcasting of a and b (right-aligned); element strides 0 on broadcast axes
  static void broadcast(const TDesc &a, const TDesc &b, std::vector<int64_t> &shape, std::vector<int64_t> &sa,
                        std::vector<int64_t> &sb) {
    const size_t r = std::max(a.rank(), b.rank());
    shape.assign(r, 1);
    sa.assign(r, 0);
    sb.assign(r, 0);
    for (size_t i = 0; i < r; ++i) {
      const long long ia = (long long)i - (long long)(r - a.rank()), ib = (long long)i - (long long)(r - b.rank());
      const size_t da = ia >= 0 ? a.shape[ia] : 1, db = ib >= 0 ? b.shape[ib] : 1;
      check(da == db || da == 1 || db == 1, "MultiplyFactory", "shapes are not broadcastable");
      shape[i] = (int64_t)std::max(da, db);
      if (ia >= 0 && da != 1) sa[i] = (int64_t)(a.strides[ia] / size_of(a.dtype));
      if (ib >= 0 && db != 1) sb[i] = (int64_t)(b.strides[ib] / size_of(b.dtype));
    }
  }
  InferResult infer(std::span<const TDesc> in, const Json &) const override {
    const char *who = "MultiplyFactory";
    check(in.size() == 2, who, "expected exactly two inputs");
    for (const auto &d : in) {
      require_device(d, who);
      check(d.dtype == DType::F32 || d.dtype == DType::CF32, who, "inputs must be F32 or CF32");
    }
    std::vector<int64_t> shape, sa, sb;
    broadcast(in[0], in[1], shape, sa, sb);
    const DType od = (in[0].dtype == DType::CF32 || in[1].dtype == DType::CF32) ? DType::CF32 : DType::F32;
    return sync_result({in[0], in[1]}, {TDesc(std::vector<size_t>(shape.begin(), shape.end()), od, in[0].mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    (void)infer(in, s);
    auto t = std::make_unique<MultiplyTask>();
    t->st = ctx.stream;
    broadcast(in[0], in[1], t->shape, t->sa, t->sb);
    t->ac = in[0].dtype == DType::CF32;
    t->bc = in[1].dtype == DType::CF32;
    return t;
  }
};

// ================================================================== FFTShiftNp
class FFTShiftTask : public ISyncTask {
 public:
  hipStream_t st;
  std::vector<int64_t> shape;
  std::vector<int32_t> mask;
  int esz;
  OpResult execute(SyncCtx &c) override {
    HFX_CHECK(hfx_fftshift(c.inputs[0].data(), c.outputs[0].data(), esz, (int32_t)shape.size(), shape.data(),
                           mask.data(), hs(st)));
    return OpResult::Ok;
  }
};

class FFTShiftFactory : public ISyncTaskFactory {
 public:
  static std::vector<int32_t> mask_of(const TDesc &d, const Json &s) {
    std::vector<int32_t> m(d.rank(), 0);
    if (!s.contains("axes") || s.at("axes").is_null()) {
      std::fill(m.begin(), m.end(), 1);
      return m;
    }
    for (const auto &a : s.at("axes").items()) {
      const int ax = norm_axis(a.integer(), d.rank(), "FFTShiftFactory");
      check(!m[ax], "FFTShiftFactory", "axes must be distinct");
      m[ax] = 1;
    }
    return m;
  }
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = "FFTShiftFactory";
    check(in.size() == 1, who, "expected exactly one input");
    require_device(in[0], who);
    check(in[0].is_contiguous(), who, "input must be contiguous");
    check(in[0].rank() >= 1, who, "input must have rank >= 1");
    (void)mask_of(in[0], s);
    return sync_result({in[0]}, {TDesc(in[0].shape, in[0].dtype, in[0].mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &s, const SyncCreateCtx &ctx) const override {
    (void)infer(in, s);
    auto t = std::make_unique<FFTShiftTask>();
    t->st = ctx.stream;
    t->shape.assign(in[0].shape.begin(), in[0].shape.end());
    t->mask = mask_of(in[0], s);
    t->esz = (int)size_of(in[0].dtype);
    return t;
  }
};

// ================================================================== DopplerMoments (fused)
hfx_plan_desc moments_desc(const Json &j, const TDesc &d) {
  hfx_plan_desc p;
  hfx_plan_desc_init(&p);
  const size_t r = d.rank();
  p.window = (int32_t)d.shape[r - 3];
  p.height = (int32_t)d.shape[r - 2];
  p.width = (int32_t)d.shape[r - 1];
  p.stride = (int32_t)j.get_int("stride", p.window);
  p.batch_lanes = 1;  // one window per execute: no second lane
  p.fs_hz = j.get("fs", 37e3);
  p.weight_fs_hz = j.get("weight_fs", 37e3);
  p.lambda_m = f32(j, "lambda");
  p.dx_m = f32(j, "dx");
  p.dy_m = j.contains("dy") ? f32(j, "dy") : p.dx_m;
  p.z_m = f32(j, "z");
  p.bin_begin = (int32_t)(j.contains("time_z_begin") ? j.at("time_z_begin").integer() : j.at("bin_begin").integer());
  p.bin_end = (int32_t)(j.contains("time_z_end") ? j.at("time_z_end").integer() : j.at("bin_end").integer());
  const std::string prop = j.contains("prop") ? j.at("prop").str() : "fresnel";
  if (prop == "fresnel")
    p.prop = HFX_PROP_FRESNEL;
  else if (prop == "angular_spectrum")
    p.prop = HFX_PROP_ANGULAR_SPECTRUM;
  else
    throw std::invalid_argument("DopplerMomentsFactory inference error: unknown prop '" + prop + "'");
  const std::string prec = j.contains("precision") ? j.at("precision").str() : "fp32";
  if (prec == "fp32")
    p.precision = HFX_PREC_FP32;
  else if (prec == "fp16")
    p.precision = HFX_PREC_FP16;
  else if (prec == "bf16")
    p.precision = HFX_PREC_BF16;
  else if (prec == "fp16_split")
    p.precision = HFX_PREC_FP16_SPLIT;
  else
    throw std::invalid_argument("DopplerMomentsFactory inference error: unknown precision '" + prec + "'");
  p.flags = (j.get_bool("fftshift", false) ? HFX_FLAG_FFTSHIFT : 0u) | (j.get_bool("power2", false) ? HFX_FLAG_POWER2 : 0u) |
            (j.get_bool("both_sidebands", false) ? HFX_FLAG_BOTH_SIDEBANDS : 0u);
  if (j.contains("padding") && !j.at("padding").is_null()) {
    p.pad_width = (int32_t)j.at("padding").at("width").integer();
    p.pad_height = (int32_t)j.at("padding").at("height").integer();
  }
  if (j.contains("filter") && !j.at("filter").is_null()) {
    const auto &f = j.at("filter");
    p.filter_enabled = 1;
    p.r_inner = (double)f.at("r_inner").integer();
    p.r_outer = (double)f.at("r_outer").integer();
    p.s_inner = (double)f.at("s_inner").integer();
    p.s_outer = (double)f.at("s_outer").integer();
  }
  p.split_bin = (int32_t)j.get_int("split_bin", 0);
  if (j.contains("flatfield") && !j.at("flatfield").is_null()) {
    const auto &ff = j.at("flatfield");
    p.flatfield_enabled = 1;
    if (ff.contains("cutoff_period_m")) {
      // sigma = 0.187 * period / image pitch, the Fresnel image pitch being lambda |z| / (N d)
      // (graph_builder.cc:32-78), in FP32 as the reference's settings hold them
      const float period = (float)ff.at("cutoff_period_m").number();
      float pitch_y = (float)p.dy_m, pitch_x = (float)p.dx_m;
      if (p.prop == HFX_PROP_FRESNEL) {
        const float lam = (float)p.lambda_m, z = std::fabs((float)p.z_m);
        if (!(z > 0.0f)) fail("DopplerMomentsFactory", "Fresnel output pitch requires non-zero propagation distance");
        pitch_y = lam * z / ((float)p.height * (float)p.dy_m);
        pitch_x = lam * z / ((float)p.width * (float)p.dx_m);
      }
      if (!(period > 0.0f && pitch_y > 0.0f && pitch_x > 0.0f))
        fail("DopplerMomentsFactory", "flatfield cutoff period and image pitches must be positive");
      p.flatfield_sigma_y = 0.187f * period / pitch_y;
      p.flatfield_sigma_x = 0.187f * period / pitch_x;
    } else {
      p.flatfield_sigma_y = (float)ff.at("sigma_y").number();
      p.flatfield_sigma_x = (float)ff.at("sigma_x").number();
    }
  }
  return p;
}

class MomentsTask : public ISyncTask {
 public:
  hipStream_t st;
  Json settings;
  TDesc idesc;
  hfx_plan *plan = nullptr;
  ~MomentsTask() override {
    if (plan) hfx_plan_destroy(plan);
  }
  OpResult execute(SyncCtx &c) override {
    const int64_t T = (int64_t)idesc.shape[idesc.rank() - 3];
    HFX_CHECK(hfx_moments(plan, reinterpret_cast<const uint16_t *>(c.inputs[0].data()), T, 0,
                          reinterpret_cast<float *>(c.outputs[0].data()), hs(st)));
    return OpResult::Ok;
  }
};

class DopplerMomentsFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &j) const override {
    const char *who = "DopplerMomentsFactory";
    check(in.size() == 1, who, "expected exactly one input");
    const TDesc &d = in[0];
    check(d.dtype == DType::U16, who, "input must be a U16 frame window");
    require_device(d, who);
    check(d.is_contiguous(), who, "input must be contiguous");
    check(d.rank() == 3 || (d.rank() == 4 && d.shape[0] == 1), who, "input must be [T, H, W] or [1, T, H, W]");
    const auto p = moments_desc(j, d);
    check(p.lambda_m > 0 && p.dx_m > 0 && p.dx_m == p.dy_m, who, "invalid optics (lambda > 0, dx == dy > 0)");
    check(p.bin_begin >= 0 && p.bin_end > p.bin_begin, who, "band must satisfy 0 <= begin < end");
    check(!p.flatfield_enabled || (p.flatfield_sigma_y > 0.0f && p.flatfield_sigma_x > 0.0f), who,
          "flatfield sigma must be positive");
    const size_t n_out = p.split_bin > 0 ? 5 : 3;
    const size_t oh = p.pad_height > 0 ? (size_t)p.pad_height : (size_t)p.height;
    const size_t ow = p.pad_width > 0 ? (size_t)p.pad_width : (size_t)p.width;
    return sync_result({d}, {TDesc({n_out, oh, ow}, DType::F32, d.mem_loc)});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &j, const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto t = std::make_unique<MomentsTask>();
    t->st = ctx.stream;
    t->settings = j;
    t->idesc = in[0];
    const auto p = moments_desc(j, in[0]);
    HFX_CHECK(hfx_plan_create(&p, hs(ctx.stream), &t->plan));
    return t;
  }
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &j,
                                    const SyncCreateCtx &ctx) const override {
    auto *t = dynamic_cast<MomentsTask *>(old.get());
    if (!t || in.size() != 1 || in[0] != t->idesc) return create(in, j, ctx);
    if (j == t->settings) {
      t->st = ctx.stream;
      return old;
    }
    Json a = j, b = t->settings;
    a["z"] = Json(0.0);
    b["z"] = Json(0.0);
    // a flatfield set by its cutoff period follows the Fresnel image pitch, which moves with z
    const bool pitch_follows_z = j.contains("flatfield") && !j.at("flatfield").is_null() &&
                                 j.at("flatfield").contains("cutoff_period_m") &&
                                 moments_desc(j, in[0]).prop == HFX_PROP_FRESNEL;
    if (a == b && !pitch_follows_z) {
      (void)infer(in, j);
      t->st = ctx.stream;
      HFX_CHECK(hfx_plan_update_z(t->plan, f32(j, "z"), hs(ctx.stream)));
      t->settings = j;
      return old;
    }
    return create(in, j, ctx);
  }
};

// ================================================================== BatchQueue (async ring)
class BatchQueue : public IAsyncTask {
 public:
  TDesc idesc, odesc;
  size_t nb_slots = 0, input_size = 0, element_size = 0, output_stride = 0;
  curaii::unique_device_ptr<std::byte> dbuf;
  curaii::unique_host_ptr<std::byte> hbuf;
  std::byte *buf = nullptr;
  std::atomic<size_t> write_idx{0}, read_idx{0};

  size_t used() const {
    const size_t w = write_idx.load(std::memory_order_acquire), r = read_idx.load(std::memory_order_acquire);
    return w >= r ? w - r : w + nb_slots - r;
  }
  std::optional<TView> acquire_input(int index) override {
    if (index != 0) throw std::out_of_range("BatchQueue::acquire_input: invalid index");
    if (nb_slots - used() <= input_size) return std::nullopt;  // a full ring keeps one slot free
    auto &st = storage_access().owned_input_storage(0);
    st.ptr = buf + write_idx.load(std::memory_order_relaxed) * element_size;
    return TView{idesc, &st};
  }
  OpResult try_push(AsyncPushCtx &) override {
    size_t w = write_idx.load(std::memory_order_relaxed) + input_size;
    if (w >= nb_slots) w = 0;
    storage_access().owned_input_storage(0).ptr = nullptr;
    write_idx.store(w, std::memory_order_release);
    return OpResult::Ok;
  }
  OpResult try_pop(AsyncPopCtx &ctx) override {
    if (used() < output_stride) return OpResult::NotReady;
    auto &st = storage_access().owned_output_storage(0);
    st.ptr = buf + read_idx.load(std::memory_order_relaxed) * element_size;
    ctx.outputs[0] = TView{odesc, &st};
    return OpResult::Ok;
  }
  void release_output(int index) override {
    if (index != 0) throw std::out_of_range("BatchQueue::release_output: invalid index");
    size_t r = read_idx.load(std::memory_order_relaxed) + output_stride;
    if (r == nb_slots) r = 0;
    storage_access().owned_output_storage(0).ptr = nullptr;
    read_idx.store(r, std::memory_order_release);
  }
};

class BatchQueueFactory : public IAsyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = "BatchQueueFactory";
    check(in.size() == 1, who, "BatchQueue task must have exactly one input");
    check(in[0].rank() > 0, who, "BatchQueue task input must have rank > 0");
    check(in[0].is_contiguous(), who, "BatchQueue task input must be contiguous");
    const long long cap = s.at("target_capacity").integer(), size = s.at("output_size").integer(),
                    stride = s.at("output_stride").integer();
    check(cap > 0, who, "BatchQueue task target capacity must be > 0");
    check(size > 0, who, "BatchQueue task output size must be > 0");
    check(stride > 0, who, "BatchQueue task output stride must be > 0");
    check(stride % size == 0, who, "BatchQueue task output stride must be a multiple of output size");
    TDesc out = in[0];
    out.shape[0] = (size_t)size;
    out.strides = TDesc::contiguous_strides(out.shape, out.dtype);
    InferResult r;
    r.input_descs = {in[0]};
    r.output_descs = {out};
    r.owned_inputs = {true};
    r.owned_outputs = {true};
    r.kind = TaskKind::Async;
    return r;
  }
  // queue geometry without its ring buffer
  std::unique_ptr<BatchQueue> make(std::span<const TDesc> in, const Json &s) const {
    const auto r = infer(in, s);
    auto q = std::make_unique<BatchQueue>();
    q->idesc = in[0];
    q->odesc = r.output_descs[0];
    q->input_size = in[0].shape[0];
    q->output_stride = (size_t)s.at("output_stride").integer();
    q->element_size = in[0].num_bytes() / in[0].shape[0];
    // smallest multiple of lcm(input batch, stride) holding capacity + one input batch
    // (batch_queue.cc:176-181, 232-239)
    const long long x = (long long)q->input_size, y = (long long)q->output_stride;
    const long long k = s.at("target_capacity").integer() + x;
    const long long base = std::lcm(x, y);
    q->nb_slots = (size_t)(base * ((k + base - 1) / base));
    return q;
  }
  std::unique_ptr<IAsyncTask> create(std::span<const TDesc> in, const Json &s, const AsyncCreateCtx &ctx) const override {
    auto q = make(in, s);
    const size_t bytes = q->nb_slots * q->element_size;
    if (in[0].mem_loc == MemLoc::Device) {
      q->dbuf = curaii::make_unique_device_ptr<std::byte>(bytes);
      q->buf = q->dbuf.get();
    } else {
      q->hbuf = curaii::make_unique_host_ptr<std::byte>(bytes);
      q->buf = q->hbuf.get();
    }
    (void)ctx;
    return q;
  }
  // A queue of the same byte size and memory location takes over the old ring (emptied);
  // anything else is recreated (batch_queue.cc:271-306).
  std::unique_ptr<IAsyncTask> update(std::unique_ptr<IAsyncTask> old, std::span<const TDesc> in, const Json &s,
                                     const AsyncCreateCtx &ctx) const override {
    auto *ob = dynamic_cast<BatchQueue *>(old.get());
    if (!ob) return create(in, s, ctx);
    auto q = make(in, s);
    if (q->nb_slots * q->element_size != ob->nb_slots * ob->element_size || in[0].mem_loc != ob->idesc.mem_loc)
      return create(in, s, ctx);
    q->dbuf = std::move(ob->dbuf);
    q->hbuf = std::move(ob->hbuf);
    q->buf = ob->buf;
    return q;
  }
};

// ================================================================== SlidingAverage (async)
// Rolling mean of the last window_size pushed frames (slide_avg.cu:52-313): pushes land in
// a ring of target_capacity + window_size slots; each push folds the new frame in and the
// frame window_size pushes older out of a running sum, and stores the running mean in that
// older frame's slot, which the consumer pops in order.  An optional host U8 input of shape
// [1] discards the push when zero; `discard_first` drops the first accepted pushes.
class SlidingAverage : public IAsyncTask {
 public:
  TDesc idesc, odesc;
  size_t ws = 1, cap = 1, discard_first = 0, discarded = 0, nb_slots = 0, element_size = 0;
  hipStream_t producer = nullptr;
  curaii::unique_device_ptr<std::byte> buf;
  curaii::unique_device_ptr<float> ravg;
  std::atomic<size_t> avg_idx{0}, write_idx{0}, read_idx{0};
  Json settings;

  size_t dist(size_t r, size_t w) const { return w >= r ? w - r : w + nb_slots - r; }
  std::optional<TView> acquire_input(int index) override {
    if (index != 0) throw std::out_of_range("SlidingAverage::acquire_input: invalid index");
    if (nb_slots - dist(read_idx.load(std::memory_order_acquire), write_idx.load()) <= 1) return std::nullopt;
    auto &st = storage_access().owned_input_storage(0);
    st.ptr = buf.get() + write_idx.load(std::memory_order_relaxed) * element_size;
    return TView{idesc, &st};
  }
  OpResult try_push(AsyncPushCtx &ctx) override {
    const bool invalid = ctx.inputs.size() == 2 && *reinterpret_cast<const uint8_t *>(ctx.inputs[1].data()) == 0;
    if (invalid || discarded < discard_first) {
      if (!invalid) ++discarded;
      storage_access().owned_input_storage(0).ptr = nullptr;
      return OpResult::Ok;
    }
    const size_t w = write_idx.load(std::memory_order_relaxed), a = avg_idx.load(std::memory_order_relaxed);
    auto *frame = reinterpret_cast<const float *>(buf.get() + w * element_size);
    auto *old = reinterpret_cast<float *>(buf.get() + a * element_size);
    HFX_CHECK(hfx_slide_avg_step(frame, old, ravg.get(), (int64_t)(element_size / 4), (int32_t)ws, hs(producer)));
    HIP_CHECK(hipStreamSynchronize(producer));
    avg_idx.store(a + 1 == nb_slots ? 0 : a + 1, std::memory_order_release);
    storage_access().owned_input_storage(0).ptr = nullptr;
    write_idx.store(w + 1 == nb_slots ? 0 : w + 1, std::memory_order_release);
    return OpResult::Ok;
  }
  OpResult try_pop(AsyncPopCtx &ctx) override {
    const size_t r = read_idx.load(std::memory_order_relaxed);
    if (dist(r, write_idx.load(std::memory_order_acquire)) <= ws) return OpResult::NotReady;
    auto &st = storage_access().owned_output_storage(0);
    st.ptr = buf.get() + r * element_size;
    ctx.outputs[0] = TView{odesc, &st};
    return OpResult::Ok;
  }
  void release_output(int index) override {
    if (index != 0) throw std::out_of_range("SlidingAverage::release_output: invalid index");
    const size_t r = read_idx.load(std::memory_order_relaxed);
    storage_access().owned_output_storage(0).ptr = nullptr;
    read_idx.store(r + 1 == nb_slots ? 0 : r + 1, std::memory_order_release);
  }
};

class SlidingAverageFactory : public IAsyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    auto check_ = [](bool c, const char *what) {
      if (!c) throw std::logic_error(what);
    };
    const long long cap = s.at("target_capacity").integer(), ws = s.at("window_size").integer();
    check_(in.size() == 1 || in.size() == 2, "expected one input and optional valid");
    const TDesc &d = in[0];
    check_(d.dtype == DType::F32, "input must be float32");
    check_(d.rank() == 3, "input must be 3D");
    check_(d.shape[0] == 1, "input batch size must be 1");
    check_(d.mem_loc == MemLoc::Device, "input must be in device memory");
    check_(d.is_contiguous(), "input must be contiguous");
    check_(cap > 0, "target_capacity must be greater or equal than zero");
    check_(ws > 0, "window_size must be greater or equal than zero");
    if (in.size() == 2) {
      check_(in[1].dtype == DType::U8, "valid input must be uint8");
      check_(in[1].mem_loc == MemLoc::Host, "valid input must be in host memory");
      check_(in[1].shape == std::vector<size_t>{1}, "valid input must have shape [1]");
    }
    InferResult r;
    r.input_descs.assign(in.begin(), in.end());
    r.output_descs = {d};
    r.owned_inputs = in.size() == 1 ? std::vector<bool>{true} : std::vector<bool>{true, false};
    r.owned_outputs = {true};
    r.kind = TaskKind::Async;
    return r;
  }
  std::unique_ptr<IAsyncTask> create(std::span<const TDesc> in, const Json &s, const AsyncCreateCtx &ctx) const override {
    const auto r = infer(in, s);
    auto t = std::make_unique<SlidingAverage>();
    t->settings = s;
    t->idesc = in[0];
    t->odesc = r.output_descs[0];
    t->cap = (size_t)s.at("target_capacity").integer();
    t->ws = (size_t)s.at("window_size").integer();
    t->discard_first = (size_t)s.get_int("discard_first", 0);
    t->nb_slots = t->cap + t->ws;
    t->element_size = in[0].num_bytes();
    t->producer = ctx.producer_stream;
    t->buf = curaii::make_unique_device_ptr<std::byte>(t->nb_slots * t->element_size);
    t->ravg = curaii::make_unique_device_ptr<float>(t->element_size / 4);
    HIP_CHECK(hipMemsetAsync(t->ravg.get(), 0, t->element_size, ctx.producer_stream));
    HIP_CHECK(hipMemsetAsync(t->buf.get(), 0, t->nb_slots * t->element_size, ctx.producer_stream));
    HIP_CHECK(hipStreamSynchronize(ctx.producer_stream));
    t->avg_idx = t->nb_slots - t->ws;
    t->write_idx = 0;
    t->read_idx = t->nb_slots - 1;
    return t;
  }
  std::unique_ptr<IAsyncTask> update(std::unique_ptr<IAsyncTask> old, std::span<const TDesc> in, const Json &s,
                                     const AsyncCreateCtx &ctx) const override {
    (void)infer(in, s);
    auto *t = dynamic_cast<SlidingAverage *>(old.get());
    if (t && t->settings == s && t->idesc == in[0]) {
      t->producer = ctx.producer_stream;
      return old;
    }
    return create(in, s, ctx);
  }
};

// ================================================================== DualReaderBatchQueue (async)
// Ring with two read cursors (dual_reader_batch_queue.cc:48-181): each pop exposes the current
// frame I[n], the delayed frame I[n - h] with h = (window_size - 1) / 2 (zero scratch until
// the delayed reader has a frame) and a host U8 flag that turns 1 once n >= window_size - 1.
// A slot is reused only after both readers released it.
class DualReaderBatchQueue : public IAsyncTask {
 public:
  TDesc idesc, odesc;
  size_t ws = 1, delay = 0, slot_count = 0, input_size = 1, element_size = 0;
  curaii::unique_device_ptr<std::byte> dbuf, dscratch;
  curaii::unique_host_ptr<std::byte> hbuf, hscratch;
  std::byte *buf = nullptr, *scratch = nullptr;
  std::atomic<size_t> write_idx{0}, delayed_idx{0};
  size_t current_idx = 0, sequence = 0;
  bool pop_active = false, delayed_active = false, cur_released = false, del_released = false;

  size_t inc(size_t i) const { return i + 1 == slot_count ? 0 : i + 1; }
  size_t dist(size_t b, size_t e) const { return e >= b ? e - b : e + slot_count - b; }
  std::optional<TView> acquire_input(int index) override {
    if (index != 0) throw std::out_of_range("DualReaderBatchQueue::acquire_input: invalid index");
    if (slot_count - dist(delayed_idx.load(std::memory_order_acquire), write_idx.load()) <= input_size)
      return std::nullopt;
    auto &st = storage_access().owned_input_storage(0);
    st.ptr = buf + write_idx.load(std::memory_order_relaxed) * element_size;
    return TView{idesc, &st};
  }
  OpResult try_push(AsyncPushCtx &) override {
    size_t w = write_idx.load(std::memory_order_relaxed) + input_size;
    if (w >= slot_count) w = 0;
    storage_access().owned_input_storage(0).ptr = nullptr;
    write_idx.store(w, std::memory_order_release);
    return OpResult::Ok;
  }
  OpResult try_pop(AsyncPopCtx &ctx) override {
    if (dist(current_idx, write_idx.load(std::memory_order_acquire)) < 1) return OpResult::NotReady;
    if (pop_active) throw std::logic_error("DualReaderBatchQueue::try_pop: prior output was not released");
    auto &cur = storage_access().owned_output_storage(0);
    auto &del = storage_access().owned_output_storage(1);
    cur.ptr = buf + current_idx * element_size;
    delayed_active = sequence >= delay;
    del.ptr = delayed_active ? buf + delayed_idx.load(std::memory_order_relaxed) * element_size : scratch;
    ctx.outputs[0] = TView{odesc, &cur};
    ctx.outputs[1] = TView{odesc, &del};
    *reinterpret_cast<uint8_t *>(ctx.outputs[2].data()) = sequence >= ws - 1 ? 1 : 0;
    cur_released = del_released = false;
    pop_active = true;
    return OpResult::Ok;
  }
  void release_output(int index) override {
    if (index != 0 && index != 1) throw std::out_of_range("DualReaderBatchQueue::release_output: invalid index");
    if (!pop_active) throw std::logic_error("DualReaderBatchQueue::release_output: no active output");
    storage_access().owned_output_storage((size_t)index).ptr = nullptr;
    (index == 0 ? cur_released : del_released) = true;
    if (!cur_released || !del_released) return;
    current_idx = inc(current_idx);
    if (delayed_active) delayed_idx.store(inc(delayed_idx.load(std::memory_order_relaxed)), std::memory_order_release);
    ++sequence;
    pop_active = false;
  }
};

class DualReaderBatchQueueFactory : public IAsyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &s) const override {
    const char *who = "DualReaderBatchQueueFactory";
    const long long cap = s.at("target_capacity").integer(), ws = s.at("window_size").integer();
    check(in.size() == 1, who, "task must have exactly one input");
    const TDesc &d = in[0];
    check(d.rank() > 0, who, "input rank must be positive");
    check(d.shape[0] > 0, who, "input leading dimension must be positive");
    check(d.is_contiguous(), who, "input must be contiguous");
    check(cap > 0, who, "target_capacity must be positive");
    check(ws > 0, who, "window_size must be positive");
    TDesc out = d;
    out.shape[0] = 1;
    InferResult r;
    r.input_descs = {d};
    r.output_descs = {out, out, TDesc({1}, DType::U8, MemLoc::Host)};
    r.owned_inputs = {true};
    r.owned_outputs = {true, true, false};
    r.kind = TaskKind::Async;
    return r;
  }
  std::unique_ptr<IAsyncTask> create(std::span<const TDesc> in, const Json &s, const AsyncCreateCtx &ctx) const override {
    const auto r = infer(in, s);
    auto q = std::make_unique<DualReaderBatchQueue>();
    q->idesc = in[0];
    q->odesc = r.output_descs[0];
    q->ws = (size_t)s.at("window_size").integer();
    q->delay = (q->ws - 1) / 2;
    q->input_size = in[0].shape[0];
    q->element_size = in[0].num_bytes() / q->input_size;
    // multiple of the input batch holding capacity + one batch + the delay + one slot
    // (dual_reader_batch_queue.cc:242-246)
    const size_t minimum = (size_t)s.at("target_capacity").integer() + q->input_size + q->delay + 1;
    q->slot_count = ((minimum + q->input_size - 1) / q->input_size) * q->input_size;
    const size_t bytes = q->slot_count * q->element_size;
    if (in[0].mem_loc == MemLoc::Host) {
      q->hbuf = curaii::make_unique_host_ptr<std::byte>(bytes);
      q->hscratch = curaii::make_unique_host_ptr<std::byte>(q->element_size);
      q->buf = q->hbuf.get();
      q->scratch = q->hscratch.get();
      std::memset(q->scratch, 0, q->element_size);
    } else {
      q->dbuf = curaii::make_unique_device_ptr<std::byte>(bytes);
      q->dscratch = curaii::make_unique_device_ptr<std::byte>(q->element_size);
      q->buf = q->dbuf.get();
      q->scratch = q->dscratch.get();
      HIP_CHECK(hipMemsetAsync(q->scratch, 0, q->element_size, ctx.consumer_stream));
      HIP_CHECK(hipStreamSynchronize(ctx.consumer_stream));
    }
    return q;
  }
};

}  // namespace

// ================================================================== Holofile (source)
// Frame source over a .holo recording (sources/holofile.cc:30-434) through libhfx's reader
// (include/hfx_holofile.h).  Settings {path, load_kind Live|CPUCached|GPUCached,
// start_frame, end_frame, batch_size, max_fps?, keep_cursor=true}; output [batch_size, H, W]
// U8 or U16 (8 or 16 bits per pixel).  Live reads each batch from the file into the
// caller's host buffer; CPUCached / GPUCached preload frames [start_frame, end_frame) into
// pinned host / device memory at create and hand out views into that cache (owned output).
// A batch that would pass end_frame wraps the cursor back to start_frame.  max_fps paces
// batches at batch_size / max_fps seconds apart.
//
// One deliberate difference: the reference's cached modes index the cache at
// frame_idx * frame_bytes although the cache starts at start_frame (holofile.cc:223 against
// :294, :332-337), which reads past the cache whenever start_frame > 0; here the view is at
// (frame_idx - start_frame) * frame_bytes, the frame the cursor names.
enum class LoadKind { Live, CPUCached, GPUCached };

struct HolofileSettings {
  std::string path;
  LoadKind load_kind = LoadKind::Live;
  long long start_frame = 0, end_frame = 0, batch_size = 0;
  std::optional<double> max_fps;
  bool keep_cursor = true;
};

HolofileSettings holofile_settings(const Json &j) {
  HolofileSettings s;
  s.path = j.at("path").str();
  const std::string &lk = j.at("load_kind").str();
  if (lk == "Live") s.load_kind = LoadKind::Live;
  else if (lk == "CPUCached") s.load_kind = LoadKind::CPUCached;
  else if (lk == "GPUCached") s.load_kind = LoadKind::GPUCached;
  else throw std::invalid_argument("Invalid LoadKind string: " + lk);
  s.start_frame = j.at("start_frame").integer();
  s.end_frame = j.at("end_frame").integer();
  s.batch_size = j.at("batch_size").integer();
  if (j.contains("max_fps") && !j.at("max_fps").is_null()) s.max_fps = (double)j.at("max_fps").integer();
  s.keep_cursor = j.get_bool("keep_cursor", true);
  return s;
}

struct HolofileCloser {
  void operator()(hfx_holofile *f) const { (void)hfx_holofile_close(f); }
};
using HolofileReader = std::unique_ptr<hfx_holofile, HolofileCloser>;

HolofileReader open_holofile(const std::string &path, hfx_holofile_info &info) {
  hfx_holofile *f = nullptr;
  if (hfx_holofile_open(path.c_str(), &f) != HFX_OK) throw std::runtime_error(hfx_last_error());
  HolofileReader r(f);
  if (hfx_holofile_get_info(f, &info) != HFX_OK) throw std::runtime_error(hfx_last_error());
  return r;
}

void holofile_call(hfx_status rc) {
  if (rc == HFX_OK) return;
  const std::string msg = hfx_last_error();
  if (rc == HFX_EOF) throw std::out_of_range(msg);
  throw HfxError(rc, msg);
}

class HolofileTask : public ISyncTask {
 public:
  using Clock = std::chrono::steady_clock;
  HolofileSettings settings;
  HolofileReader reader;
  hfx_holofile_info header{};
  long long frame_idx = 0;
  TDesc odesc;
  std::optional<Clock::time_point> next_batch_start;
  std::byte *buf = nullptr;  // the cache of frames [start_frame, end_frame), cached kinds only
  curaii::unique_host_ptr<std::byte> h_buf;
  curaii::unique_device_ptr<std::byte> d_buf;

  // Fixed-rate pacing (holofile.cc:153-183): wait for the batch's slot, then advance the
  // slot by one period, skipping periods already in the past.
  bool pace(SyncCtx &ctx) {
    if (!settings.max_fps) {
      next_batch_start.reset();
      return true;
    }
    auto period = std::chrono::duration_cast<Clock::duration>(
        std::chrono::duration<double>((double)settings.batch_size / *settings.max_fps));
    if (period <= Clock::duration::zero()) period = Clock::duration{1};
    if (!next_batch_start) next_batch_start = Clock::now();
    while (Clock::now() < *next_batch_start) {
      if (ctx.cancelled && ctx.cancelled->load(std::memory_order_relaxed)) return false;
      std::this_thread::sleep_for(std::min<Clock::duration>(*next_batch_start - Clock::now(), std::chrono::milliseconds(1)));
    }
    do {
      *next_batch_start += period;
    } while (*next_batch_start < Clock::now());
    return true;
  }

  void release_output(int index) override {
    if (settings.load_kind == LoadKind::Live) throw std::logic_error("Cannot release output for Live load kind");
    if (index != 0) throw std::out_of_range("Holofile task has only one output at index 0");
  }

  OpResult execute(SyncCtx &ctx) override {
    if (!pace(ctx)) return OpResult::Cancelled;
    if (frame_idx + settings.batch_size > settings.end_frame) {
      if (settings.load_kind == LoadKind::Live) holofile_call(hfx_holofile_seek(reader.get(), settings.start_frame));
      frame_idx = settings.start_frame;
    }
    if (settings.load_kind == LoadKind::Live) {
      holofile_call(hfx_holofile_read(reader.get(), ctx.outputs[0].data(), settings.batch_size));
    } else {
      auto &st = storage_access().owned_output_storage(0);
      st = Storage{odesc.mem_loc, odesc.num_bytes(),
                   buf + (size_t)(frame_idx - settings.start_frame) * (size_t)header.frame_bytes};
      ctx.outputs[0] = TView{odesc, &st};
    }
    frame_idx += settings.batch_size;
    return OpResult::Ok;
  }
};

class HolofileFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &j) const override {
    const char *who = "HolofileFactory";
    const auto s = holofile_settings(j);
    hfx_holofile_info h{};
    (void)open_holofile(s.path, h);
    check(in.empty(), who, "Holofile task must have no inputs");
    check(s.start_frame >= 0, who, "Invalid frame range");
    check(s.start_frame < s.end_frame, who, "Invalid frame range");
    check(s.batch_size > 0, who, "Batch size must be positive");
    check(!s.max_fps || *s.max_fps > 0, who, "max_fps must be positive when provided");
    check(s.end_frame <= (long long)h.frame_count, who, "end_frame exceeds total frames in file");
    check(s.batch_size <= s.end_frame - s.start_frame, who, "Batch size exceeds available frames");
    check(h.bits_per_pixel == 8 || h.bits_per_pixel == 16, who,
          "Unsupported bits_per_pixel: " + std::to_string(h.bits_per_pixel));
    TDesc out({(size_t)s.batch_size, (size_t)h.height, (size_t)h.width}, h.bits_per_pixel == 8 ? DType::U8 : DType::U16,
              s.load_kind == LoadKind::GPUCached ? MemLoc::Device : MemLoc::Host);
    InferResult r = sync_result({}, {out});
    r.owned_outputs = {s.load_kind != LoadKind::Live};
    return r;
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &j, const SyncCreateCtx &ctx) const override {
    const auto r = infer(in, j);
    auto t = std::make_unique<HolofileTask>();
    t->settings = holofile_settings(j);
    t->reader = open_holofile(t->settings.path, t->header);
    holofile_call(hfx_holofile_seek(t->reader.get(), t->settings.start_frame));
    const long long frames = t->settings.end_frame - t->settings.start_frame;
    const size_t bytes = (size_t)frames * (size_t)t->header.frame_bytes;
    if (t->settings.load_kind == LoadKind::CPUCached) {
      t->h_buf = curaii::make_unique_host_ptr<std::byte>(bytes);
      t->buf = t->h_buf.get();
      holofile_call(hfx_holofile_read(t->reader.get(), t->buf, frames));
    } else if (t->settings.load_kind == LoadKind::GPUCached) {
      t->d_buf = curaii::make_unique_device_ptr<std::byte>(bytes);
      t->buf = t->d_buf.get();
      holofile_call(hfx_holofile_read_to_device(t->reader.get(), t->buf, frames, hs(ctx.stream)));
    }
    t->frame_idx = t->settings.start_frame;
    t->odesc = r.output_descs[0];
    return t;
  }
  // Same file and load kind (and, for the cached kinds, the same frame range) keep the
  // reader and the cache; keep_cursor keeps the cursor when it stays inside the range
  // (holofile.cc:361-431).
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &j,
                                    const SyncCreateCtx &ctx) const override {
    auto *o = dynamic_cast<HolofileTask *>(old.get());
    if (!o) return create(in, j, ctx);
    const auto r = infer(in, j);
    const auto s = holofile_settings(j);
    const bool live = s.load_kind == LoadKind::Live;
    const bool bounds_same = o->settings.start_frame == s.start_frame && o->settings.end_frame == s.end_frame;
    const bool pace_same = o->settings.batch_size == s.batch_size && o->settings.max_fps == s.max_fps;
    if (!(o->settings.path == s.path && o->settings.load_kind == s.load_kind && (live || bounds_same)))
      return create(in, j, ctx);
    auto t = std::make_unique<HolofileTask>();
    t->settings = s;
    t->reader = std::move(o->reader);
    t->header = o->header;
    t->buf = o->buf;
    t->h_buf = std::move(o->h_buf);
    t->d_buf = std::move(o->d_buf);
    t->frame_idx = s.start_frame;
    if (s.keep_cursor) {
      t->frame_idx = o->frame_idx;
      if (t->frame_idx < s.start_frame || t->frame_idx >= s.end_frame) t->frame_idx = s.start_frame;
    }
    if (live && t->frame_idx != o->frame_idx) holofile_call(hfx_holofile_seek(t->reader.get(), t->frame_idx));
    t->odesc = r.output_descs[0];
    if (pace_same) t->next_batch_start = o->next_batch_start;
    return t;
  }
};

// ================================================================== Memcpy
// Host/device transfer (memcpy.cc:105-196): the output is the input's descriptor in the
// target memory; device transfers are enqueued on the node's stream, host to host copies
// run on the calling thread.
class MemcpyTask : public ISyncTask {
 public:
  hipStream_t st;
  MemLoc target = MemLoc::Device;
  OpResult execute(SyncCtx &c) override {
    const TView &in = c.inputs[0];
    const size_t n = c.outputs[0].desc.num_bytes();
    const MemLoc src = in.desc.mem_loc;
    if (src == MemLoc::Host && target == MemLoc::Host) {
      std::memcpy(c.outputs[0].data(), in.data(), n);
    } else {
      const hipMemcpyKind kind = src == MemLoc::Host ? hipMemcpyHostToDevice
                                 : target == MemLoc::Host ? hipMemcpyDeviceToHost
                                                          : hipMemcpyDeviceToDevice;
      HIP_CHECK(hipMemcpyAsync(c.outputs[0].data(), in.data(), n, kind, st));
    }
    return OpResult::Ok;
  }
};

MemLoc memcpy_target(const Json &j) {
  const std::string &t = j.at("target").str();
  if (t == "Host") return MemLoc::Host;
  if (t == "Device") return MemLoc::Device;
  throw std::invalid_argument("Invalid Target string: " + t);
}

class MemcpyFactory : public ISyncTaskFactory {
 public:
  InferResult infer(std::span<const TDesc> in, const Json &j) const override {
    const MemLoc target = memcpy_target(j);
    check(in.size() == 1, "MemcpyFactory", "Memcpy task must have exactly one input");
    check(in[0].is_contiguous(), "MemcpyFactory", "Memcpy task input must be contiguous");
    TDesc out = in[0];
    out.mem_loc = target;
    return sync_result({in[0]}, {out});
  }
  std::unique_ptr<ISyncTask> create(std::span<const TDesc> in, const Json &j, const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    auto t = std::make_unique<MemcpyTask>();
    t->st = ctx.stream;
    t->target = memcpy_target(j);
    return t;
  }
  std::unique_ptr<ISyncTask> update(std::unique_ptr<ISyncTask> old, std::span<const TDesc> in, const Json &j,
                                    const SyncCreateCtx &ctx) const override {
    (void)infer(in, j);
    if (auto *t = dynamic_cast<MemcpyTask *>(old.get()); t && t->target == memcpy_target(j)) {
      t->st = ctx.stream;
      return old;
    }
    return create(in, j, ctx);
  }
};

void register_hip_nodes(Registry &r) {
  r.register_async("BatchQueue", std::make_unique<BatchQueueFactory>());
  r.register_async("SlidingAverage", std::make_unique<SlidingAverageFactory>());
  r.register_async("DualReaderBatchQueue", std::make_unique<DualReaderBatchQueueFactory>());
  r.register_sync("Conversion", std::make_unique<ConversionFactory>());
  r.register_sync("Reshape", std::make_unique<ReshapeFactory>());
  r.register_sync("Slice", std::make_unique<SliceFactory>());
  r.register_sync("Copy", std::make_unique<CopyFactory>());
  r.register_sync("Concatenate", std::make_unique<ConcatenateFactory>());
  r.register_sync("RFFT", std::make_unique<DftFactory<true>>());
  r.register_sync("FFT", std::make_unique<DftFactory<false>>());
  r.register_sync("FresnelDiffraction", std::make_unique<FresnelDiffractionFactory>());
  r.register_sync("AngularSpectrum", std::make_unique<AngularSpectrumFactory>());
  r.register_sync("Filter2D", std::make_unique<Filter2DFactory>());
  r.register_sync("Flatfield", std::make_unique<FlatfieldFactory>());
  r.register_sync("MeanAbs", std::make_unique<ReduceFactory<true>>());
  r.register_sync("Mean", std::make_unique<ReduceFactory<false>>());
  r.register_sync("Abs", std::make_unique<AbsFactory>());
  r.register_sync("Arange", std::make_unique<ArangeFactory>());
  r.register_sync("Multiply", std::make_unique<MultiplyFactory>());
  r.register_sync("FFTShiftNp", std::make_unique<FFTShiftFactory>());
  r.register_sync("DopplerMoments", std::make_unique<DopplerMomentsFactory>());
  r.register_sync("Holofile", std::make_unique<HolofileFactory>());
  r.register_sync("Memcpy", std::make_unique<MemcpyFactory>());
}

}  // namespace holoflow::core

// ====================================================================== node-runner ABI
using namespace holoflow::core;

struct hfx_node : public IOStorageAccess {
  std::string kind;
  Json settings;
  std::vector<TDesc> in_descs;
  InferResult info;
  hipStream_t stream = nullptr;
  std::unique_ptr<ISyncTask> sync;
  std::unique_ptr<IAsyncTask> async;
  Storage owned_in[4] = {}, owned_out[4] = {};
  Storage &owned_input_storage(size_t i) override {
    if (i >= 4) throw std::out_of_range("owned input index");
    return owned_in[i];
  }
  Storage &owned_output_storage(size_t i) override {
    if (i >= 4) throw std::out_of_range("owned output index");
    return owned_out[i];
  }
};

namespace {

const Registry &registry() {
  static Registry *r = [] {
    auto *reg = new Registry();
    register_hip_nodes(*reg);
    return reg;
  }();
  return *r;
}

std::vector<TDesc> parse_descs(const char *text) {
  std::vector<TDesc> v;
  const Json j = Json::parse(text ? text : "[]");
  for (const auto &d : j.items()) v.push_back(tdesc_from_json(d));
  return v;
}

template <typename F> hfx_status guarded(F &&f) {
  try {
    f();
    hfx::set_last_error("");
    return HFX_OK;
  } catch (const std::logic_error &e) {  // invalid_argument, out_of_range, logic_error
    hfx::set_last_error(e.what());
    return HFX_E_INVALID;
  } catch (const curaii::HfxError &e) {
    hfx::set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    hfx::set_last_error(e.what());
    return HFX_E_HIP;
  }
}

hfx_status copy_out(const std::string &s, char *buf, int64_t cap) {
  if (!buf || cap <= (int64_t)s.size()) {
    hfx::set_last_error("output buffer too small (" + std::to_string(s.size() + 1) + " bytes needed)");
    return HFX_E_INVALID;
  }
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return HFX_OK;
}

Json infer_json(const InferResult &r) {
  Json j = Json::object();
  j["kind"] = Json(r.kind == TaskKind::Sync ? "Sync" : "Async");
  Json outs = Json::array();
  for (const auto &d : r.output_descs) outs.push_back(to_json(d));
  j["output_descs"] = outs;
  Json ip = Json::array();
  for (const auto &p : r.in_place) ip.push_back(Json(std::vector<int>{p.in_idx, p.out_idx}));
  j["in_place"] = ip;
  j["owned_inputs"] = Json(std::vector<bool>(r.owned_inputs.begin(), r.owned_inputs.end()));
  j["owned_outputs"] = Json(std::vector<bool>(r.owned_outputs.begin(), r.owned_outputs.end()));
  return j;
}

void make_task(hfx_node *n) {
  if (const auto *sf = registry().sync_factory(n->kind)) {
    n->sync = sf->create(n->in_descs, n->settings, SyncCreateCtx{n->stream});
    n->sync->bind_storage_access(n);
  } else if (const auto *af = registry().async_factory(n->kind)) {
    n->async = af->create(n->in_descs, n->settings, AsyncCreateCtx{n->stream, n->stream});
    n->async->bind_storage_access(n);
  } else {
    throw std::invalid_argument("unknown node kind: " + n->kind);
  }
}

}  // namespace

extern "C" {

hfx_status hfx_node_kinds(char *buf, int64_t cap) {
  std::string s;
  hfx_status rc = guarded([&] { s = Json(registry().kinds()).dump(); });
  return rc != HFX_OK ? rc : copy_out(s, buf, cap);
}

hfx_status hfx_node_infer(const char *kind, const char *settings_json, const char *input_descs_json, char *buf,
                          int64_t cap) {
  std::string s;
  hfx_status rc = guarded([&] {
    const auto descs = parse_descs(input_descs_json);
    const Json settings = Json::parse(settings_json ? settings_json : "{}");
    s = infer_json(registry().factory(kind ? kind : "").infer(descs, settings)).dump();
  });
  return rc != HFX_OK ? rc : copy_out(s, buf, cap);
}

hfx_status hfx_node_create(const char *kind, const char *settings_json, const char *input_descs_json,
                           hfx_stream_t stream, hfx_node **out) {
  if (!out) return HFX_E_INVALID;
  *out = nullptr;
  return guarded([&] {
    auto n = std::make_unique<hfx_node>();
    n->kind = kind ? kind : "";
    n->settings = Json::parse(settings_json ? settings_json : "{}");
    n->in_descs = parse_descs(input_descs_json);
    n->stream = reinterpret_cast<hipStream_t>(stream);
    n->info = registry().factory(n->kind).infer(n->in_descs, n->settings);
    make_task(n.get());
    *out = n.release();
  });
}

hfx_status hfx_node_update(hfx_node *n, const char *settings_json, const char *input_descs_json, hfx_stream_t stream,
                           int32_t *reused) {
  if (!n) return HFX_E_INVALID;
  return guarded([&] {
    const Json settings = Json::parse(settings_json ? settings_json : "{}");
    const auto descs = parse_descs(input_descs_json);
    const auto info = registry().factory(n->kind).infer(descs, settings);
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    bool kept = false;
    if (const auto *sf = registry().sync_factory(n->kind)) {
      ISyncTask *before = n->sync.get();
      n->sync = sf->update(std::move(n->sync), descs, settings, SyncCreateCtx{st});
      n->sync->bind_storage_access(n);
      kept = n->sync.get() == before;
    } else if (const auto *af = registry().async_factory(n->kind)) {
      IAsyncTask *before = n->async.get();
      n->async = af->update(std::move(n->async), descs, settings, AsyncCreateCtx{st, st});
      n->async->bind_storage_access(n);
      kept = n->async.get() == before;
    }
    n->settings = settings;
    n->in_descs = descs;
    n->info = info;
    n->stream = st;
    if (reused) *reused = kept ? 1 : 0;
  });
}

hfx_status hfx_node_execute(hfx_node *n, void *const *inputs, int32_t n_inputs, void **outputs,
                            int32_t n_outputs) {
  if (!n || !n->sync) {
    hfx::set_last_error("not a synchronous node");
    return HFX_E_INVALID;
  }
  return guarded([&] {
    if ((size_t)n_inputs != n->in_descs.size() || (size_t)n_outputs != n->info.output_descs.size())
      throw std::invalid_argument("execute: wrong number of inputs or outputs");
    std::vector<Storage> st(n_inputs + n_outputs);
    std::vector<TView> in(n_inputs), out(n_outputs);
    for (int i = 0; i < n_inputs; ++i) {
      st[i] = Storage{n->in_descs[i].mem_loc, n->in_descs[i].num_bytes(), static_cast<std::byte *>(inputs[i])};
      in[i] = TView{n->in_descs[i], &st[i]};
    }
    const auto &od = n->info.output_descs;
    for (int i = 0; i < n_outputs; ++i) {
      // an output the task owns is a view it hands back; otherwise caller storage
      const bool owned = i < (int)n->info.owned_outputs.size() && n->info.owned_outputs[i];
      st[n_inputs + i] = Storage{od[i].mem_loc, od[i].num_bytes(), owned ? nullptr : static_cast<std::byte *>(outputs[i])};
      out[i] = TView{od[i], &st[n_inputs + i]};
    }
    std::atomic<bool> cancelled{false};
    SyncCtx ctx{in, out, &cancelled};
    if (n->sync->execute(ctx) != OpResult::Ok) throw std::runtime_error("execute did not return Ok");
    for (int i = 0; i < n_outputs; ++i)
      if (i < (int)n->info.owned_outputs.size() && n->info.owned_outputs[i]) outputs[i] = out[i].data();
  });
}

hfx_status hfx_node_acquire_input(hfx_node *n, void **slot) {
  if (!n || !n->async || !slot) return HFX_E_INVALID;
  hfx_status rc = HFX_OK;
  hfx_status g = guarded([&] {
    auto v = n->async->acquire_input(0);
    if (!v) {
      rc = HFX_NOT_READY;
      *slot = nullptr;
    } else {
      *slot = v->data();
    }
  });
  return g != HFX_OK ? g : rc;
}

hfx_status hfx_node_push(hfx_node *n, void *const *extra_inputs, int32_t n_extra) {
  if (!n || !n->async || n_extra < 0 || (n_extra > 0 && !extra_inputs)) return HFX_E_INVALID;
  return guarded([&] {
    if ((size_t)n_extra + 1 != n->in_descs.size())
      throw std::invalid_argument("push: expected " + std::to_string(n->in_descs.size() - 1) + " extra inputs");
    std::vector<Storage> st(n->in_descs.size());
    std::vector<TView> views(n->in_descs.size());
    views[0] = TView{n->in_descs[0], &n->owned_in[0]};
    for (int i = 0; i < n_extra; ++i) {
      const TDesc &d = n->in_descs[i + 1];
      st[i + 1] = Storage{d.mem_loc, d.num_bytes(), static_cast<std::byte *>(extra_inputs[i])};
      views[i + 1] = TView{d, &st[i + 1]};
    }
    std::atomic<bool> cancelled{false};
    AsyncPushCtx ctx{std::span<TView>(views), &cancelled};
    (void)n->async->try_push(ctx);
  });
}

hfx_status hfx_node_pop(hfx_node *n, void **outputs, int32_t n_outputs) {
  if (!n || !n->async || !outputs) return HFX_E_INVALID;
  hfx_status rc = HFX_OK;
  hfx_status g = guarded([&] {
    const auto &od = n->info.output_descs;
    if ((size_t)n_outputs != od.size())
      throw std::invalid_argument("pop: expected " + std::to_string(od.size()) + " outputs");
    std::vector<Storage> st(od.size());
    std::vector<TView> views(od.size());
    for (size_t i = 0; i < od.size(); ++i) {
      // outputs the task does not own are caller storage the task writes into
      if (!n->info.owned_outputs[i]) {
        st[i] = Storage{od[i].mem_loc, od[i].num_bytes(), static_cast<std::byte *>(outputs[i])};
        views[i] = TView{od[i], &st[i]};
      }
    }
    std::atomic<bool> cancelled{false};
    AsyncPopCtx ctx{std::span<TView>(views), &cancelled};
    if (n->async->try_pop(ctx) != OpResult::Ok) {
      rc = HFX_NOT_READY;
      return;
    }
    for (size_t i = 0; i < od.size(); ++i)
      if (n->info.owned_outputs[i]) outputs[i] = views[i].data();
  });
  return g != HFX_OK ? g : rc;
}

hfx_status hfx_node_release_output(hfx_node *n, int32_t index) {
  if (!n || (!n->async && !n->sync)) return HFX_E_INVALID;
  return guarded([&] {
    if (n->async) n->async->release_output(index);
    else n->sync->release_output(index);
  });
}

hfx_status hfx_node_destroy(hfx_node *n) {
  delete n;
  return HFX_OK;
}

}  // extern "C"
