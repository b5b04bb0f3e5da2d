/*
 * hfx_holofile.h - the .holo recording format in front of the Doppler-moment path.
 *
 * Layout (src/holofile/include/holofile/holofile.hh:87-110, src/holofile.cc:44-260): a
 * packed 64-byte little-endian header {u32 magic 0x4F4C4F48 ("HOLO"), u16 version (66),
 * u16 bits per pixel (8 or 16), u32 width, u32 height, u32 frame count, u64 data size in
 * bytes, u8 endianness, 35 bytes padding}, then the raw frames, then an optional JSON
 * footer (the pipeline settings, at most 1 MiB, starting with '{').
 *
 * The reader validates the magic number and the frame size, and reads the footer when one
 * is present (a missing or malformed footer is not an error: footer text is then empty).
 * Errors: HFX_E_INVALID for a malformed file (message as the reference's exceptions:
 * "Holofile: Invalid magic number", ...), HFX_EOF when fewer frames remain than asked.
 */
#ifndef HFX_HOLOFILE_H_
#define HFX_HOLOFILE_H_

#include <stdint.h>

#include "hfx.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef HFX_EOF
#define HFX_EOF 6
#endif

typedef struct hfx_holofile_info {
  uint32_t width, height;
  uint16_t bits_per_pixel; /* 8 or 16 */
  uint16_t version;
  uint32_t frame_count;
  uint64_t data_size;      /* bytes of frame data after the header */
  uint64_t frame_bytes;    /* width * height * bits_per_pixel / 8 */
  int32_t has_footer;
} hfx_holofile_info;

typedef struct hfx_holofile hfx_holofile;

hfx_status hfx_holofile_open(const char *path, hfx_holofile **out);
hfx_status hfx_holofile_close(hfx_holofile *f);
hfx_status hfx_holofile_get_info(const hfx_holofile *f, hfx_holofile_info *info);
/* Footer JSON text (NUL-terminated; "" without a footer). */
hfx_status hfx_holofile_footer(const hfx_holofile *f, char *buf, int64_t cap);
/* Reader::seek / tell (holofile.cc:165-180): the next read starts at frame `frame`. */
hfx_status hfx_holofile_seek(hfx_holofile *f, int64_t frame);
int64_t hfx_holofile_tell(const hfx_holofile *f);
/* Reader::read_frames (holofile.cc:182-201) into host memory. */
hfx_status hfx_holofile_read(hfx_holofile *f, void *host_dst, int64_t n_frames);
/* The same frames into device memory, staged through the reader's pinned buffers
 * (double-buffered: the file read of one chunk overlaps the copy of the previous one);
 * returns after the copies on `stream` completed. */
hfx_status hfx_holofile_read_to_device(hfx_holofile *f, void *dev_dst, int64_t n_frames, hfx_stream_t stream);
/* Writer (holofile.cc:203-290): header, n_frames frames from host memory, optional footer
 * JSON (NULL or "" for none). */
hfx_status hfx_holofile_write(const char *path, uint32_t width, uint32_t height, uint16_t bits_per_pixel,
                              const void *host_frames, int64_t n_frames, const char *footer_json);

#ifdef __cplusplus
}
#endif

#endif /* HFX_HOLOFILE_H_ */
