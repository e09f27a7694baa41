// vvclf_recon.hip — the steps either side of the in-loop chain:
//
//   vvclf_lmcs_forward_luma     LMCS forward mapping of inter-CU luma
//                               rectangles (src/ctu.rs:1701-1725,
//                               src/lmcs.rs:186-195), through a 64 K-entry
//                               table built on the host
//   vvclf_lmcs_chroma_residual  chroma residual scaling per chroma TB
//                               (src/ctu.rs:1752-1815, :734-742,
//                               src/lmcs.rs:243-279): one wave per block
//                               reduces the neighbouring luma, picks the
//                               piece and folds the residual into the
//                               prediction
//   vvclf_conformance_crop,     output crop (src/stream.rs:816-842) and the
//   vvclf_picture_md5,          per-plane MD5 of the conformance convention
//   vvclf_job_output_md5        (tests/jvet_conformance.rs:400-430); for a
//                               job the crop is packed on the device
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

namespace {

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

// Forward LUT entry for samples the reference cannot map (>= 2^BitDepth
// below 16 bits): never a mapped value, since those are <= 2^BitDepth - 1.
constexpr uint16_t kFwdBad = 0xffffu;

// One wave per rectangle.  Rectangles of a power-of-two width up to 64 are
// walked 64 / w rows at a time (every lane busy); others row by row.
__global__ void __launch_bounds__(64) k_lmcs_fwd(uint16_t* y, uint32_t pitch, uint32_t W, uint32_t H,
                                                 const vvclf_rect* rects, const uint16_t* lut, int bd16,
                                                 uint32_t* bad) {
  const vvclf_rect r = rects[blockIdx.x];
  const uint32_t lane = threadIdx.x;
  const uint32_t x0 = r.x, y0 = r.y;
  const uint32_t x1 = min(r.x + r.w, W), y1 = min(r.y + r.h, H);
  if (x0 >= x1 || y0 >= y1) return;
  const uint32_t w = x1 - x0;
  uint32_t hit = 0;
  if (w <= 64 && (w & (w - 1)) == 0) {
    const uint32_t lw = __builtin_ctz(w), rows = 64u >> lw;
    const uint32_t c = lane & (w - 1), r0 = lane >> lw;
    for (uint32_t yy = y0 + r0; yy < y1; yy += rows) {
      uint16_t* p = y + (size_t)yy * pitch + x0 + c;
      const uint16_t m = lut[*p];
      if (m == kFwdBad && !bd16) hit = 1;
      else *p = m;
    }
  } else {
    for (uint32_t yy = y0; yy < y1; yy++)
      for (uint32_t xx = x0 + lane; xx < x1; xx += 64) {
        uint16_t* p = y + (size_t)yy * pitch + xx;
        const uint16_t m = lut[*p];
        if (m == kFwdBad && !bd16) hit = 1;
        else *p = m;
      }
  }
  if (hit) atomicOr(bad, 1u);
}

struct crs_tables {
  uint32_t pivot[17];  // LmcsPivot
  uint32_t csc[16];    // ChromaScaleCoeff
  uint32_t min_bin, max_bin, bd;
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wave per chroma TB.  The luma plane is only read.
__global__ void __launch_bounds__(64) k_chroma_crs(const uint16_t* luma, uint32_t lp, uint32_t W, uint32_t H,
                                                   uint16_t* cb, uint16_t* cr, uint32_t cp,
                                                   const vvclf_crs_block* blocks, const int32_t* res,
                                                   crs_tables t) {
  const vvclf_crs_block b = blocks[blockIdx.x];
  const uint32_t lane = threadIdx.x;
  // eq. 1216: left column / top row of sizeY mapped-domain luma samples
  uint32_t s = 0;
  if (lane < b.size_y) {
    if (b.avail_l) s += luma[(size_t)min(b.y_cu + lane, H - 1) * lp + (b.x_cu - 1)];
    if (b.avail_t) s += luma[(size_t)(b.y_cu - 1) * lp + min(b.x_cu + lane, W - 1)];
  }
  const uint32_t sum = wave_sum(s);
  const uint32_t cnt = (b.avail_l ? b.size_y : 0u) + (b.avail_t ? b.size_y : 0u);
  // eq. 1216 (shift by the trailing zeros of cnt) / eq. 1217
  const uint32_t avg = cnt ? (sum + (cnt >> 1)) >> __builtin_ctz(cnt) : 1u << (t.bd - 1);
  uint32_t idx = t.min_bin;  // idxYInv, src/lmcs.rs:209-220
  while (idx <= t.max_bin && avg >= t.pivot[idx + 1]) idx++;
  const uint32_t vs = t.csc[min(idx, 15u)];
  // eqs. 1219 / 1220 and Clip1(pred + res)
  const int32_t lo = -(1 << t.bd), hi = (1 << t.bd) - 1;
  uint16_t* pl = b.c_idx == 1 ? cb : cr;
  const int32_t* rb = res + b.res_offset;
  const uint32_t n = b.w * b.h;
  for (uint32_t i = lane; i < n; i += 64) {
    const uint32_t ry = i / b.w, rx = i - ry * b.w;
    const int32_t rv = min(max(rb[i], lo), hi);
    const uint32_t mag = (uint32_t)(((uint64_t)(uint32_t)abs(rv) * vs + 1024u) >> 11);
    uint16_t* p = pl + (size_t)(b.y + ry) * cp + b.x + rx;
    const int32_t v = (int32_t)*p + (rv < 0 ? -(int32_t)mag : (int32_t)mag);
    *p = (uint16_t)min(max(v, 0), hi);
  }
}

// Crop of one plane packed as the conformance hash reads it: one byte per
// sample (the low byte) at 8 bits, two bytes little-endian above.
__global__ void __launch_bounds__(256) k_pack_crop(const uint16_t* p, uint32_t pitch, uint32_t x0, uint32_t y0,
                                                   uint32_t w, uint32_t h, int wide, uint8_t* out) {
  const uint32_t yy = blockIdx.y;
  if (yy >= h) return;
  const uint16_t* row = p + (size_t)(y0 + yy) * pitch + x0;
  for (uint32_t x = blockIdx.x * 256 + threadIdx.x; x < w; x += gridDim.x * 256) {
    const uint16_t v = row[x];
    if (wide) {
      out[((size_t)yy * w + x) * 2] = (uint8_t)(v & 0xff);
      out[((size_t)yy * w + x) * 2 + 1] = (uint8_t)(v >> 8);
    } else {
      out[(size_t)yy * w + x] = (uint8_t)(v & 0xff);
    }
  }
}

// ---------------------------------------------------------------------------
// MD5 (RFC 1321), host side
// ---------------------------------------------------------------------------
struct md5_ctx {
  uint32_t a = 0x67452301u, b = 0xefcdab89u, c = 0x98badcfeu, d = 0x10325476u;
  uint64_t len = 0;
  uint8_t buf[64];
  size_t fill = 0;

  static const uint32_t* table() {  // T[i] = floor(2^32 |sin(i + 1)|)
    static uint32_t t[64];
    static bool init = [] {
      for (int i = 0; i < 64; i++) t[i] = (uint32_t)(std::fabs(std::sin((double)(i + 1))) * 4294967296.0);
      return true;
    }();
    (void)init;
    return t;
  }
  static uint32_t rotl(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }
  void block(const uint8_t* p) {
    static const int S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
    const uint32_t* T = table();
    uint32_t X[16];
    for (int i = 0; i < 16; i++)
      X[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
             ((uint32_t)p[4 * i + 3] << 24);
    uint32_t A = a, B = b, C = c, D = d;
    for (int i = 0; i < 64; i++) {
      uint32_t f;
      int g;
      if (i < 16) {
        f = (B & C) | (~B & D);
        g = i;
      } else if (i < 32) {
        f = (B & D) | (C & ~D);
        g = (5 * i + 1) & 15;
      } else if (i < 48) {
        f = B ^ C ^ D;
        g = (3 * i + 5) & 15;
      } else {
        f = C ^ (B | ~D);
        g = (7 * i) & 15;
      }
      const uint32_t nb = B + rotl(A + f + T[i] + X[g], S[i >> 4][i & 3]);
      A = D;
      D = C;
      C = B;
      B = nb;
    }
    a += A;
    b += B;
    c += C;
    d += D;
  }
  void update(const uint8_t* p, size_t n) {
    len += n;
    if (fill) {
      const size_t k = std::min(n, 64 - fill);
      memcpy(buf + fill, p, k);
      fill += k;
      p += k;
      n -= k;
      if (fill < 64) return;
      block(buf);
      fill = 0;
    }
    for (; n >= 64; p += 64, n -= 64) block(p);
    memcpy(buf, p, n);
    fill = n;
  }
  void finish(uint8_t out[16]) {
    const uint64_t bits = len * 8;
    const uint8_t one = 0x80, zero = 0;
    update(&one, 1);
    while (fill != 56) update(&zero, 1);
    uint8_t l[8];
    for (int i = 0; i < 8; i++) l[i] = (uint8_t)(bits >> (8 * i));
    update(l, 8);
    const uint32_t w[4] = {a, b, c, d};
    for (int i = 0; i < 16; i++) out[i] = (uint8_t)(w[i / 4] >> (8 * (i % 4)));
  }
};

void md5_plane(const uint16_t* s, size_t stride, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, bool wide,
               uint8_t out[16]) {
  md5_ctx m;
  std::vector<uint8_t> row((size_t)w * 2);
  for (uint32_t y = y0; y < y0 + h; y++) {
    const uint16_t* r = s + (size_t)y * stride + x0;
    if (wide) {
      for (uint32_t x = 0; x < w; x++) {
        row[2 * x] = (uint8_t)(r[x] & 0xff);
        row[2 * x + 1] = (uint8_t)(r[x] >> 8);
      }
      m.update(row.data(), (size_t)w * 2);
    } else {
      for (uint32_t x = 0; x < w; x++) row[x] = (uint8_t)(r[x] & 0xff);
      m.update(row.data(), w);
    }
  }
  m.finish(out);
}

// Plane crops of a 4:2:0 / 4:0:0 picture (luma, then chroma = luma / 2).
int crop_rects(uint32_t W, uint32_t H, uint32_t Wc, uint32_t Hc, uint32_t cfi, const vvclf_crop& c,
               vvclf_crop out[3]) {
  if ((uint64_t)c.x + c.w > W || (uint64_t)c.y + c.h > H)
    return fail(VVCLF_ERR_INVALID, "crop %ux%u at (%u,%u) outside the %ux%u picture", c.w, c.h, c.x, c.y, W, H);
  out[0] = c;
  if (cfi != 0) {
    const vvclf_crop cc{c.x / 2, c.y / 2, c.w / 2, c.h / 2};
    if ((uint64_t)cc.x + cc.w > Wc || (uint64_t)cc.y + cc.h > Hc)
      return fail(VVCLF_ERR_INVALID, "chroma crop outside the %ux%u chroma planes", Wc, Hc);
    out[1] = out[2] = cc;
  }
  return VVCLF_OK;
}

// A device buffer from the drop-in context's stream-ordered pool.
struct dbuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  hipError_t alloc(size_t n, hipStream_t st) {
    s = st;
    return hipMallocAsync(&p, std::max<size_t>(n, 4), st);
  }
  ~dbuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
};

}  // namespace

extern "C" {

int vvclf_lmcs_forward_luma(vvclf_plane* luma, const vvclf_lmcs_binding* b, const vvclf_rect* rects,
                            uint32_t n_rects) {
  clear_error();
  if (!luma || !b) return fail(VVCLF_ERR_INVALID, "lmcs_forward_luma: NULL argument");
  if (!b->derived || n_rects == 0) return VVCLF_OK;
  if (!rects) return fail(VVCLF_ERR_INVALID, "lmcs_forward_luma: NULL rects");
  const vvclf_lmcs_derived& d = *b->derived;
  if (d.bit_depth < 8 || d.bit_depth > 16) return fail(VVCLF_ERR_INVALID, "lmcs: bit depth %u", d.bit_depth);
  if (!luma->samples || luma->stride < luma->width) return fail(VVCLF_ERR_INVALID, "lmcs_forward_luma: bad plane");
  const uint32_t W = luma->width, H = luma->height;
  if (W == 0 || H == 0) return VVCLF_OK;
  // the rectangles, clipped to the plane, must not overlap (each sample is
  // mapped once, as by the reference's per-CU calls)
  std::vector<uint8_t> seen((size_t)W * H, 0);
  for (uint32_t i = 0; i < n_rects; i++) {
    const vvclf_rect& r = rects[i];
    const uint32_t x1 = (uint32_t)std::min<uint64_t>((uint64_t)r.x + r.w, W),
                   y1 = (uint32_t)std::min<uint64_t>((uint64_t)r.y + r.h, H);
    for (uint32_t y = r.y; y < y1; y++)
      for (uint32_t x = r.x; x < x1; x++) {
        uint8_t& v = seen[(size_t)y * W + x];
        if (v) return fail(VVCLF_ERR_INVALID, "lmcs_forward_luma: rectangle %u overlaps another at (%u,%u)", i, x, y);
        v = 1;
      }
  }
  // forward LUT over every u16 value: Clip3(0, 2^BD - 1, eq. 1213)
  std::vector<uint16_t> lut(65536);
  const int64_t hi = ((int64_t)1 << d.bit_depth) - 1;
  const uint32_t lg = d.bit_depth - 4;
  for (uint32_t s = 0; s < 65536; s++) {
    const uint32_t idx = s >> lg;
    if (idx >= 16) {
      lut[s] = kFwdBad;
      continue;
    }
    const int64_t prod = (int64_t)d.scale_coeff[idx] * ((int64_t)s - (int64_t)d.input_pivot[idx]) + (1 << 10);
    const int64_t v = (int64_t)d.lmcs_pivot[idx] + (prod >> 11);
    lut[s] = (uint16_t)std::min<int64_t>(std::max<int64_t>(v, 0), hi);
  }
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  dbuf dy, dl, dr, dbad;
  hipError_t e = dy.alloc((size_t)W * H * 2, st);
  if (e == hipSuccess) e = dl.alloc(65536 * 2, st);
  if (e == hipSuccess) e = dr.alloc((size_t)n_rects * sizeof(vvclf_rect), st);
  if (e == hipSuccess) e = dbad.alloc(4, st);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "lmcs_forward_luma: device buffers: %s", hipGetErrorString(e));
  e = copy_rows(dy.p, (size_t)W * 2, luma->samples, (size_t)luma->stride * 2, (size_t)W * 2, H,
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dl.p, lut.data(), 65536 * 2, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dr.p, rects, (size_t)n_rects * sizeof(vvclf_rect), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(dbad.p, 0, 4, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_lmcs_fwd, dim3(n_rects), dim3(64), 0, st, (uint16_t*)dy.p, W, W, H,
                       (const vvclf_rect*)dr.p, (const uint16_t*)dl.p, d.bit_depth >= 16 ? 1 : 0,
                       (uint32_t*)dbad.p);
    e = hipGetLastError();
  }
  uint32_t bad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad.p, 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "lmcs_forward_luma");
  if (bad)
    return fail(VVCLF_ERR_INVALID, "lmcs_forward_luma: a sample of 2^%u or more lies in a rectangle", d.bit_depth);
  e = copy_rows(luma->samples, (size_t)luma->stride * 2, dy.p, (size_t)W * 2, (size_t)W * 2, H,
                       hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? VVCLF_OK : hip_fail(e, "lmcs_forward_luma: download");
}

int vvclf_lmcs_chroma_residual(vvclf_picture* pic, const vvclf_lmcs_binding* b, const vvclf_crs_block* blocks,
                               uint32_t n, const int32_t* residuals, uint64_t n_res) {
  clear_error();
  if (!pic || !b) return fail(VVCLF_ERR_INVALID, "lmcs_chroma_residual: NULL argument");
  if (!b->derived || n == 0) return VVCLF_OK;
  if (!blocks || !residuals) return fail(VVCLF_ERR_INVALID, "lmcs_chroma_residual: NULL blocks or residuals");
  const vvclf_lmcs_derived& d = *b->derived;
  if (d.bit_depth < 8 || d.bit_depth > 16 || b->max_bin_idx > 15)
    return fail(VVCLF_ERR_INVALID, "lmcs: invalid derived tables");
  const vvclf_plane &L = pic->luma, &Pb = pic->cb, &Pr = pic->cr;
  const uint32_t W = L.width, H = L.height, Wc = Pb.width, Hc = Pb.height;
  if (!L.samples || !Pb.samples || !Pr.samples || Pr.width != Wc || Pr.height != Hc || W == 0 || H == 0)
    return fail(VVCLF_ERR_INVALID, "lmcs_chroma_residual: picture planes");
  std::vector<uint8_t> seen((size_t)2 * Wc * Hc, 0);
  for (uint32_t k = 0; k < n; k++) {
    const vvclf_crs_block& q = blocks[k];
    if ((q.c_idx != 1 && q.c_idx != 2) || (uint64_t)q.x + q.w > Wc || (uint64_t)q.y + q.h > Hc ||
        q.res_offset + (uint64_t)q.w * q.h > n_res || q.size_y == 0 || q.size_y > 64 ||
        (q.avail_l && (q.x_cu == 0 || q.x_cu - 1 >= W)) || (q.avail_t && (q.y_cu == 0 || q.y_cu - 1 >= H)))
      return fail(VVCLF_ERR_INVALID, "lmcs_chroma_residual: block %u outside its planes or residuals", k);
    for (uint32_t y = q.y; y < q.y + q.h; y++)
      for (uint32_t x = q.x; x < q.x + q.w; x++) {
        uint8_t& v = seen[(size_t)(q.c_idx - 1) * Wc * Hc + (size_t)y * Wc + x];
        if (v) return fail(VVCLF_ERR_INVALID, "lmcs_chroma_residual: block %u overlaps another", k);
        v = 1;
      }
  }
  crs_tables t;
  memcpy(t.pivot, d.lmcs_pivot, sizeof(t.pivot));
  memcpy(t.csc, d.chroma_scale_coeff, sizeof(t.csc));
  t.min_bin = b->min_bin_idx;
  t.max_bin = b->max_bin_idx;
  t.bd = d.bit_depth;
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  dbuf dl, dcb, dcr, db, dres;
  hipError_t e = dl.alloc((size_t)W * H * 2, st);
  if (e == hipSuccess) e = dcb.alloc((size_t)Wc * Hc * 2, st);
  if (e == hipSuccess) e = dcr.alloc((size_t)Wc * Hc * 2, st);
  if (e == hipSuccess) e = db.alloc((size_t)n * sizeof(vvclf_crs_block), st);
  if (e == hipSuccess) e = dres.alloc((size_t)n_res * 4, st);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "lmcs_chroma_residual: device buffers: %s", hipGetErrorString(e));
  auto up = [&](void* dst, const vvclf_plane& p, uint32_t w, uint32_t h) {
    return copy_rows(dst, (size_t)w * 2, p.samples, (size_t)p.stride * 2, (size_t)w * 2, h,
                            hipMemcpyHostToDevice, st);
  };
  e = up(dl.p, L, W, H);
  if (e == hipSuccess && Wc && Hc) e = up(dcb.p, Pb, Wc, Hc);
  if (e == hipSuccess && Wc && Hc) e = up(dcr.p, Pr, Wc, Hc);
  if (e == hipSuccess) e = hipMemcpyAsync(db.p, blocks, (size_t)n * sizeof(vvclf_crs_block), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && n_res) e = hipMemcpyAsync(dres.p, residuals, (size_t)n_res * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_chroma_crs, dim3(n), dim3(64), 0, st, (const uint16_t*)dl.p, W, W, H, (uint16_t*)dcb.p,
                       (uint16_t*)dcr.p, Wc, (const vvclf_crs_block*)db.p, (const int32_t*)dres.p, t);
    e = hipGetLastError();
  }
  auto down = [&](const vvclf_plane& p, const void* src, uint32_t w, uint32_t h) {
    return copy_rows(p.samples, (size_t)p.stride * 2, src, (size_t)w * 2, (size_t)w * 2, h,
                            hipMemcpyDeviceToHost, st);
  };
  if (e == hipSuccess && Wc && Hc) e = down(Pb, dcb.p, Wc, Hc);
  if (e == hipSuccess && Wc && Hc) e = down(Pr, dcr.p, Wc, Hc);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? VVCLF_OK : hip_fail(e, "lmcs_chroma_residual");
}

int vvclf_conformance_crop(uint32_t w, uint32_t h, uint32_t cfi, const uint32_t* win, vvclf_crop* out) {
  clear_error();
  if (!out) return fail(VVCLF_ERR_INVALID, "conformance_crop: NULL out");
  const uint32_t sw = (cfi == 1 || cfi == 2) ? 2 : 1, sh = cfi == 1 ? 2 : 1;
  uint64_t l = 0, r = 0, t = 0, bo = 0;
  if (win) {
    l = (uint64_t)win[0] * sw;
    r = (uint64_t)win[1] * sw;
    t = (uint64_t)win[2] * sh;
    bo = (uint64_t)win[3] * sh;
  }
  if (l + r > w || t + bo > h) return fail(VVCLF_ERR_INVALID, "conformance window larger than the picture");
  *out = vvclf_crop{(uint32_t)l, (uint32_t)t, (uint32_t)(w - l - r), (uint32_t)(h - t - bo)};
  return VVCLF_OK;
}

int vvclf_picture_md5(const vvclf_picture* pic, uint32_t cfi, uint32_t bd, const vvclf_crop* crop,
                      uint8_t md5[3][16]) {
  clear_error();
  if (!pic || !crop || !md5) return fail(VVCLF_ERR_INVALID, "picture_md5: NULL argument");
  vvclf_crop c[3];
  int rc = crop_rects(pic->luma.width, pic->luma.height, pic->cb.width, pic->cb.height, cfi, *crop, c);
  if (rc) return rc;
  const vvclf_plane* pl[3] = {&pic->luma, &pic->cb, &pic->cr};
  for (int k = 0; k < (cfi ? 3 : 1); k++) {
    if (!pl[k]->samples && c[k].w && c[k].h) return fail(VVCLF_ERR_INVALID, "picture_md5: NULL plane");
    md5_plane(pl[k]->samples, pl[k]->stride, c[k].x, c[k].y, c[k].w, c[k].h, bd > 8, md5[k]);
  }
  return VVCLF_OK;
}

int vvclf_job_output_md5(vvclf_job* j, uint32_t bd, const vvclf_crop* crop, uint8_t md5[3][16]) {
  clear_error();
  if (!j || !crop || !md5) return fail(VVCLF_ERR_INVALID, "job_output_md5: NULL argument");
  const uint32_t cfi = j->wc ? 1 : 0;
  vvclf_crop c[3];
  int rc = crop_rects(j->w, j->h, j->wc, j->hc, cfi, *crop, c);
  if (rc) return rc;
  hipSetDevice(j->ctx->device);
  hipStream_t st = j->ctx->stream;
  hipError_t e = hipDeviceSynchronize();  // the job's work may be on any stream
  const int wide = bd > 8 ? 1 : 0;
  size_t bytes[3] = {0, 0, 0}, total = 0;
  for (int k = 0; k < (cfi ? 3 : 1); k++) total += bytes[k] = (size_t)c[k].w * c[k].h * (wide ? 2 : 1);
  dbuf dout;
  if (e == hipSuccess) e = dout.alloc(total, st);
  std::vector<uint8_t> host(total);
  size_t off = 0;
  for (int k = 0; k < (cfi ? 3 : 1) && e == hipSuccess; k++) {
    if (c[k].w && c[k].h) {
      const uint32_t pitch = k ? (uint32_t)j->d.pc : (uint32_t)j->d.py;
      hipLaunchKernelGGL(k_pack_crop, dim3(std::min<uint32_t>((c[k].w + 255) / 256, 8), c[k].h), dim3(256), 0, st,
                         job_final(j, k), pitch, c[k].x, c[k].y, c[k].w, c[k].h, wide, (uint8_t*)dout.p + off);
      e = hipGetLastError();
    }
    off += bytes[k];
  }
  if (e == hipSuccess && total) e = hipMemcpyAsync(host.data(), dout.p, total, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "job_output_md5");
  off = 0;
  for (int k = 0; k < (cfi ? 3 : 1); k++) {
    md5_ctx m;
    m.update(host.data() + off, bytes[k]);
    m.finish(md5[k]);
    off += bytes[k];
  }
  return VVCLF_OK;
}

}  // extern "C"
