// .holo reader / writer (include/hfx_holofile.h).  Host C++; device uploads go through two
// pinned staging buffers so that reading chunk i+1 from the file overlaps copying chunk i.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <algorithm>
#include <string>

#include "hfx.h"
#include "hfx_holofile.h"
#include "holoflow/core/json.hh"

namespace hfx {
void set_last_error(const std::string &msg);  // libhfx's thread-local last error (api.hip)
}

namespace {

hfx_status fail(hfx_status code, const std::string &msg) {
  hfx::set_last_error(msg);
  return code;
}

constexpr uint32_t kMagic = 0x4F4C4F48u;  // "HOLO" read as a little-endian u32
constexpr uint16_t kVersion = 66;
constexpr size_t kHeaderBytes = 64;
constexpr size_t kMaxFooter = 1u << 20;
constexpr size_t kStageBytes = 32u << 20;  // per pinned staging buffer

#pragma pack(push, 1)
struct DiskHeader {
  uint32_t magic_number;
  uint16_t version;
  uint16_t bits_per_pixel;
  uint32_t frame_width;
  uint32_t frame_height;
  uint32_t frame_count;
  uint64_t data_size_in_bytes;
  uint8_t endianness;
  char padding[35];
};
#pragma pack(pop)
static_assert(sizeof(DiskHeader) == kHeaderBytes, "holofile header is 64 bytes");

struct FileCloser {
  void operator()(FILE *f) const { fclose(f); }
};

}  // namespace

struct hfx_holofile {
  std::unique_ptr<FILE, FileCloser> file;
  DiskHeader h{};
  uint64_t frame_bytes = 0;
  int64_t cursor = 0;
  std::string footer;
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t staged[2] = {nullptr, nullptr};
  ~hfx_holofile() {
    for (int i = 0; i < 2; ++i) {
      if (staged[i]) (void)hipEventDestroy(staged[i]);
      if (stage[i]) (void)hipHostFree(stage[i]);
    }
  }
};

extern "C" {

hfx_status hfx_holofile_open(const char *path, hfx_holofile **out) {
  if (!path || !out) return fail(HFX_E_INVALID, "null argument");
  *out = nullptr;
  auto f = std::make_unique<hfx_holofile>();
  f->file.reset(fopen(path, "rb"));
  if (!f->file) return fail(HFX_E_INVALID, std::string("Failed to open \"") + path + "\"");
  if (fread(&f->h, sizeof(DiskHeader), 1, f->file.get()) != 1)
    return fail(HFX_E_INVALID, "Holofile: Incomplete header");
  if (f->h.magic_number != kMagic) return fail(HFX_E_INVALID, "Holofile: Invalid magic number");
  const uint64_t bits = (uint64_t)f->h.frame_width * f->h.frame_height * f->h.bits_per_pixel;
  if (bits % 8 != 0 || bits == 0) return fail(HFX_E_INVALID, "Holofile: Invalid frame size");
  f->frame_bytes = bits / 8;
  // footer: everything after the data section, when it is a JSON object of at most 1 MiB
  // (holofile.cc:97-163)
  const uint64_t footer_off = kHeaderBytes + f->h.data_size_in_bytes;
  if (fseeko(f->file.get(), 0, SEEK_END) == 0) {
    const int64_t size = (int64_t)ftello(f->file.get());
    if (size > (int64_t)footer_off && (uint64_t)size - footer_off <= kMaxFooter) {
      std::string text((size_t)((uint64_t)size - footer_off), '\0');
      if (fseeko(f->file.get(), (off_t)footer_off, SEEK_SET) == 0 &&
          fread(text.data(), 1, text.size(), f->file.get()) == text.size() && !text.empty() && text[0] == '{') {
        try {  // a footer that does not parse is dropped, as a missing one (holofile.cc:142-157)
          (void)holoflow::core::Json::parse(text);
          f->footer = std::move(text);
        } catch (const std::exception &) {
        }
      }
    }
  }
  if (fseeko(f->file.get(), (off_t)kHeaderBytes, SEEK_SET) != 0) return fail(HFX_E_INVALID, "Failed to seek");
  *out = f.release();
  return HFX_OK;
}

hfx_status hfx_holofile_close(hfx_holofile *f) {
  delete f;
  return HFX_OK;
}

hfx_status hfx_holofile_get_info(const hfx_holofile *f, hfx_holofile_info *info) {
  if (!f || !info) return fail(HFX_E_INVALID, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->width = f->h.frame_width;
  info->height = f->h.frame_height;
  info->bits_per_pixel = f->h.bits_per_pixel;
  info->version = f->h.version;
  info->frame_count = f->h.frame_count;
  info->data_size = f->h.data_size_in_bytes;
  info->frame_bytes = f->frame_bytes;
  info->has_footer = f->footer.empty() ? 0 : 1;
  return HFX_OK;
}

hfx_status hfx_holofile_footer(const hfx_holofile *f, char *buf, int64_t cap) {
  if (!f || !buf || cap <= (int64_t)f->footer.size()) return fail(HFX_E_INVALID, "footer buffer too small");
  std::memcpy(buf, f->footer.c_str(), f->footer.size() + 1);
  return HFX_OK;
}

hfx_status hfx_holofile_seek(hfx_holofile *f, int64_t frame) {
  if (!f || frame < 0) return fail(HFX_E_INVALID, "invalid seek");
  const uint64_t off = kHeaderBytes + (uint64_t)frame * f->frame_bytes;
  if (fseeko(f->file.get(), (off_t)off, SEEK_SET) != 0) return fail(HFX_E_INVALID, "Failed to seek");
  f->cursor = frame;
  return HFX_OK;
}

int64_t hfx_holofile_tell(const hfx_holofile *f) { return f ? f->cursor : -1; }

hfx_status hfx_holofile_read(hfx_holofile *f, void *host_dst, int64_t n_frames) {
  if (!f || (!host_dst && n_frames > 0) || n_frames < 0) return fail(HFX_E_INVALID, "invalid read");
  const size_t got = fread(host_dst, (size_t)f->frame_bytes, (size_t)n_frames, f->file.get());
  f->cursor += (int64_t)got;
  if (got != (size_t)n_frames) {
    if (ferror(f->file.get())) return fail(HFX_E_INVALID, "Failed to read frames");
    return fail(HFX_EOF, "Holofile: End of file");
  }
  return HFX_OK;
}

hfx_status hfx_holofile_read_to_device(hfx_holofile *f, void *dev_dst, int64_t n_frames, hfx_stream_t stream) {
  if (!f || (!dev_dst && n_frames > 0) || n_frames < 0) return fail(HFX_E_INVALID, "invalid read");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (int i = 0; i < 2; ++i) {
    if (!f->stage[i]) {
      if (hipHostMalloc(&f->stage[i], kStageBytes, hipHostMallocDefault) != hipSuccess)
        return fail(HFX_E_NOMEM, "pinned staging allocation failed");
      if (hipEventCreateWithFlags(&f->staged[i], hipEventDisableTiming) != hipSuccess)
        return fail(HFX_E_HIP, "event creation failed");
    }
  }
  const int64_t per = std::max<int64_t>(1, (int64_t)(kStageBytes / f->frame_bytes));
  if ((uint64_t)per * f->frame_bytes > kStageBytes) return fail(HFX_E_UNSUPPORTED, "frame larger than the staging buffer");
  int64_t done = 0;
  int b = 0;
  hfx_status rc = HFX_OK;
  while (done < n_frames) {
    const int64_t n = std::min(per, n_frames - done);
    // the buffer is free once the copy recorded on it finished
    if (hipEventSynchronize(f->staged[b]) != hipSuccess) return fail(HFX_E_HIP, "staging wait failed");
    rc = hfx_holofile_read(f, f->stage[b], n);
    if (rc != HFX_OK) break;
    if (hipMemcpyAsync(static_cast<char *>(dev_dst) + (size_t)done * f->frame_bytes, f->stage[b],
                       (size_t)n * f->frame_bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventRecord(f->staged[b], st) != hipSuccess)
      return fail(HFX_E_HIP, "upload failed");
    done += n;
    b ^= 1;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(HFX_E_HIP, "upload synchronisation failed");
  return rc;
}

hfx_status hfx_holofile_write(const char *path, uint32_t width, uint32_t height, uint16_t bits_per_pixel,
                              const void *host_frames, int64_t n_frames, const char *footer_json) {
  if (!path || (!host_frames && n_frames > 0) || n_frames < 0) return fail(HFX_E_INVALID, "null argument");
  if (bits_per_pixel != 8 && bits_per_pixel != 16) return fail(HFX_E_INVALID, "bits_per_pixel must be 8 or 16");
  DiskHeader h{};
  h.magic_number = kMagic;
  h.version = kVersion;
  h.bits_per_pixel = bits_per_pixel;
  h.frame_width = width;
  h.frame_height = height;
  h.frame_count = (uint32_t)n_frames;
  const uint64_t frame_bytes = (uint64_t)width * height * bits_per_pixel / 8;
  h.data_size_in_bytes = frame_bytes * (uint64_t)n_frames;
  h.endianness = 0;
  std::unique_ptr<FILE, FileCloser> file(fopen(path, "wb"));
  if (!file) return fail(HFX_E_INVALID, std::string("Failed to open \"") + path + "\"");
  if (fwrite(&h, sizeof(h), 1, file.get()) != 1 ||
      (n_frames > 0 && fwrite(host_frames, (size_t)frame_bytes, (size_t)n_frames, file.get()) != (size_t)n_frames))
    return fail(HFX_E_INVALID, "Failed to write frames");
  if (footer_json && footer_json[0]) {
    const size_t len = std::strlen(footer_json);
    if (fwrite(footer_json, 1, len, file.get()) != len) return fail(HFX_E_INVALID, "Failed to write footer JSON");
  }
  if (fflush(file.get()) != 0) return fail(HFX_E_INVALID, "Failed to flush file");
  return HFX_OK;
}

}  // extern "C"
