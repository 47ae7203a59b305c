// colfile.cpp -- the reference's column-file layout.
//
// A column file is a 4-byte little-endian header followed by the raw
// little-endian values.  The SSB writer stores the element size in BYTES
// (ssb-dbgen-lite.c:192-201), the synthetic writer the width in BITS
// (synth/onegen_zipf.py:73-84).  Readers ignore the header for sizing and
// derive N from the file size (ssb/cpuref.c:8-27, synth/cpuref.c:10-35);
// the synthetic loader derives the width as bytes*8/N_fk (synth/cpuref.c:68).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mbx_internal.hpp"

using namespace mbx;

extern "C" {

enum { MBX_HDR_BYTES = 0, MBX_HDR_BITS = 1 };

int mbx_colfile_write(const char *path, const void *data, size_t n,
                      uint32_t elem_bytes, int header_kind) {
  if (!path || (!data && n) ||
      (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 &&
       elem_bytes != 8))
    return fail(MBX_ERR_INVALID, "colfile_write: bad argument");
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(MBX_ERR_INVALID, std::string("cannot open ") + path);
  const uint32_t hdr = header_kind == MBX_HDR_BITS ? elem_bytes * 8 : elem_bytes;
  const unsigned char h[4] = {(unsigned char)hdr, (unsigned char)(hdr >> 8),
                              (unsigned char)(hdr >> 16),
                              (unsigned char)(hdr >> 24)};
  bool ok = std::fwrite(h, 1, 4, f) == 4;
  if (ok && n) ok = std::fwrite(data, elem_bytes, n, f) == n;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? MBX_OK : fail(MBX_ERR_INVALID, "colfile_write: short write");
}

// *data is malloc'd (free with mbx_colfile_free); *nbytes excludes the header.
int mbx_colfile_read(const char *path, void **data, size_t *nbytes,
                     uint32_t *header) {
  if (!path || !data || !nbytes)
    return fail(MBX_ERR_INVALID, "colfile_read: null");
  *data = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(MBX_ERR_INVALID, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::rewind(f);
  if (sz < 4) {
    std::fclose(f);
    return fail(MBX_ERR_INVALID, "colfile_read: file shorter than header");
  }
  unsigned char h[4];
  if (std::fread(h, 1, 4, f) != 4) {
    std::fclose(f);
    return fail(MBX_ERR_INVALID, "colfile_read: header");
  }
  if (header)
    *header = (uint32_t)h[0] | ((uint32_t)h[1] << 8) | ((uint32_t)h[2] << 16) |
              ((uint32_t)h[3] << 24);
  const size_t body = (size_t)sz - 4;
  void *p = std::malloc(body ? body : 1);
  if (!p) {
    std::fclose(f);
    return fail(MBX_ERR_OOM, "colfile_read: host allocation");
  }
  if (body && std::fread(p, 1, body, f) != body) {
    std::free(p);
    std::fclose(f);
    return fail(MBX_ERR_INVALID, "colfile_read: short read");
  }
  std::fclose(f);
  *data = p;
  *nbytes = body;
  return MBX_OK;
}

void mbx_colfile_free(void *p) { std::free(p); }

// Reads a column file straight into pool memory on the device.
int mbx_colfile_load(mbx_ctx *ctx, const char *path, void **dptr,
                     size_t *nbytes, uint32_t *header) {
  if (!ctx || !dptr) return fail(MBX_ERR_INVALID, "colfile_load: null");
  void *h = nullptr;
  size_t nb = 0;
  int st = mbx_colfile_read(path, &h, &nb, header);
  if (st) return st;
  st = mbx_alloc(ctx, nb ? nb : 1, dptr);
  if (!st) st = mbx_memcpy_h2d(ctx, *dptr, h, nb);
  std::free(h);
  if (st && *dptr) {
    mbx_free(ctx, *dptr);
    *dptr = nullptr;
  }
  if (!st && nbytes) *nbytes = nb;
  return st;
}

}  // extern "C"
