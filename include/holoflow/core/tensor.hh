// Tensor descriptors and views of the holoflow dataflow API, HIP-typed.
// Same types and meaning as the reference (src/holoflow/include/holoflow/core/tensor.hh:
// DType :29-34, MemLoc :46-50, Storage :57-61, TDesc :64-86 with byte strides and a byte
// offset, non-owning TView :91-98), so node code written against the reference reads the
// same here.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "holoflow/core/json.hh"

namespace holoflow::core {

enum class DType : uint8_t { U8, U16, F32, CF32 };
enum class MemLoc : uint8_t { Host, Device };

inline size_t size_of(DType d) noexcept {
  switch (d) {
    case DType::U8: return 1;
    case DType::U16: return 2;
    case DType::F32: return 4;
    case DType::CF32: return 8;
  }
  return 0;
}

inline std::string_view to_string(DType d) noexcept {
  switch (d) {
    case DType::U8: return "U8";
    case DType::U16: return "U16";
    case DType::F32: return "F32";
    case DType::CF32: return "CF32";
  }
  return "?";
}

inline DType dtype_from_string(const std::string &s) {
  if (s == "U8") return DType::U8;
  if (s == "U16") return DType::U16;
  if (s == "F32") return DType::F32;
  if (s == "CF32") return DType::CF32;
  throw std::invalid_argument("unknown dtype '" + s + "'");
}

struct Storage {
  MemLoc mem_loc;
  size_t bytes;
  std::byte *ptr;
};

struct TDesc {
  std::vector<size_t> shape;
  DType dtype = DType::F32;
  MemLoc mem_loc = MemLoc::Device;
  std::vector<size_t> strides;  // bytes; contiguous when constructed without strides
  size_t offset = 0;            // bytes from the start of the storage

  TDesc() = default;
  TDesc(std::vector<size_t> s, DType d, MemLoc m) : shape(std::move(s)), dtype(d), mem_loc(m) {
    strides = contiguous_strides(shape, dtype);
  }
  TDesc(std::vector<size_t> s, DType d, MemLoc m, std::vector<size_t> st, size_t off = 0)
      : shape(std::move(s)), dtype(d), mem_loc(m), strides(std::move(st)), offset(off) {}

  static std::vector<size_t> contiguous_strides(const std::vector<size_t> &shape, DType d) {
    std::vector<size_t> st(shape.size());
    size_t acc = size_of(d);
    for (size_t i = shape.size(); i-- > 0;) {
      st[i] = acc;
      acc *= shape[i];
    }
    return st;
  }
  size_t rank() const noexcept { return shape.size(); }
  size_t num_elements() const {
    size_t n = 1;
    for (size_t s : shape) n *= s;
    return n;
  }
  size_t num_bytes() const { return num_elements() * size_of(dtype); }
  bool is_contiguous() const { return strides == contiguous_strides(shape, dtype); }
  bool operator==(const TDesc &o) const {
    return shape == o.shape && dtype == o.dtype && mem_loc == o.mem_loc && strides == o.strides &&
           offset == o.offset;
  }
  bool operator!=(const TDesc &o) const { return !(*this == o); }
};

inline Json to_json(const TDesc &d) {
  Json j = Json::object();
  j["shape"] = Json(d.shape);
  j["dtype"] = Json(std::string(to_string(d.dtype)));
  j["mem_loc"] = Json(d.mem_loc == MemLoc::Device ? "Device" : "Host");
  j["strides"] = Json(d.strides);
  j["offset"] = Json(d.offset);
  return j;
}

inline TDesc tdesc_from_json(const Json &j) {
  std::vector<size_t> shape;
  for (const auto &v : j.at("shape").items()) shape.push_back((size_t)v.integer());
  const DType dt = dtype_from_string(j.at("dtype").str());
  const MemLoc ml = (j.contains("mem_loc") && j.at("mem_loc").str() == "Host") ? MemLoc::Host : MemLoc::Device;
  if (j.contains("strides")) {
    std::vector<size_t> st;
    for (const auto &v : j.at("strides").items()) st.push_back((size_t)v.integer());
    return TDesc(shape, dt, ml, st, (size_t)j.get_int("offset", 0));
  }
  TDesc d(shape, dt, ml);
  d.offset = (size_t)j.get_int("offset", 0);
  return d;
}

struct TView {
  TDesc desc = {};
  Storage *storage = nullptr;
  std::byte *data() const { return storage ? storage->ptr + desc.offset : nullptr; }
  bool is_nullptr() const { return data() == nullptr; }
};

}  // namespace holoflow::core
