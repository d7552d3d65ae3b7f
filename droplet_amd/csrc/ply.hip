// PLY ingestion straight into a device-resident scene (gfx950).
//
// Replaces loadPlySplats (src/io/ply.ts:120-239): the header is parsed on the
// host with the reference's rules (parseHeader, ply.ts:40-91), the binary
// vertex section goes to HBM in one copy as it lies in the file, and one
// kernel de-interleaves it into the scene's struct-of-arrays with the
// loader's per-splat arithmetic (Y flip, SH_C0 scaling in f64, quaternion
// normalisation by Math.hypot in f64, f_rest copy). splatBounds
// (ply.ts:242-255) is an ordered min/max reduction over the resident
// positions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dl_common.h"
#include "dl_internal.h"
#include "render.h"

namespace dl {

namespace {

constexpr double kShC0 = 0.28209479177387814;  // ply.ts:23
constexpr size_t kHeaderMax = 65536;           // ply.ts:45
constexpr uint32_t kPlyBlock = 256;
constexpr uint32_t kPlyLdsFloats = 16384;      // 64 KB of staged rows per workgroup

// ---------------------------------------------------------------- header --
// WHATWG UTF-8 decode (TextDecoder, ply.ts:43), with U+FFFD for each maximal
// invalid subpart; an ASCII byte is never consumed by an invalid sequence.
std::vector<uint32_t> utf8_decode(const uint8_t *b, size_t n) {
  std::vector<uint32_t> out;
  out.reserve(n);
  size_t i = 0;
  while (i < n) {
    const uint32_t c = b[i];
    if (c < 0x80) {
      out.push_back(c);
      i++;
      continue;
    }
    int need;
    uint32_t cp, lo = 0x80, hi = 0xBF;
    if (c >= 0xC2 && c <= 0xDF) {
      need = 1;
      cp = c & 0x1F;
    } else if (c >= 0xE0 && c <= 0xEF) {
      need = 2;
      cp = c & 0x0F;
      if (c == 0xE0) lo = 0xA0;
      if (c == 0xED) hi = 0x9F;
    } else if (c >= 0xF0 && c <= 0xF4) {
      need = 3;
      cp = c & 0x07;
      if (c == 0xF0) lo = 0x90;
      if (c == 0xF4) hi = 0x8F;
    } else {
      out.push_back(0xFFFD);
      i++;
      continue;
    }
    size_t j = i + 1;
    bool ok = true;
    for (int k = 0; k < need; k++, j++) {
      if (j >= n) {
        ok = false;
        break;
      }
      const uint32_t d = b[j];
      const uint32_t l = k == 0 ? lo : 0x80, h = k == 0 ? hi : 0xBF;
      if (d < l || d > h) {
        ok = false;
        break;
      }
      cp = (cp << 6) | (d & 0x3F);
    }
    out.push_back(ok ? cp : 0xFFFD);
    i = j;
  }
  return out;
}

// ECMAScript WhiteSpace and LineTerminator (String.prototype.trim, \s)
bool js_space(uint32_t c) {
  switch (c) {
    case 0x09: case 0x0A: case 0x0B: case 0x0C: case 0x0D: case 0x20: case 0xA0: case 0x1680:
    case 0x2028: case 0x2029: case 0x202F: case 0x205F: case 0x3000: case 0xFEFF:
      return true;
    default:
      return c >= 0x2000 && c <= 0x200A;
  }
}

using U32s = std::vector<uint32_t>;

U32s js_trim(const U32s &s) {
  size_t a = 0, b = s.size();
  while (a < b && js_space(s[a])) a++;
  while (b > a && js_space(s[b - 1])) b--;
  return U32s(s.begin() + a, s.begin() + b);
}

bool starts_with(const U32s &s, const char *p) {
  size_t i = 0;
  for (; p[i]; i++)
    if (i >= s.size() || s[i] != (uint8_t)p[i]) return false;
  return true;
}

bool contains(const U32s &s, const char *p) {
  const size_t m = strlen(p);
  for (size_t i = 0; i + m <= s.size(); i++) {
    size_t k = 0;
    while (k < m && s[i + k] == (uint8_t)p[k]) k++;
    if (k == m) return true;
  }
  return false;
}

// UTF-8 text of decoded code points (for names and messages)
std::string to_utf8(const U32s &s) {
  std::string o;
  for (uint32_t c : s) {
    if (c < 0x80) {
      o += (char)c;
    } else if (c < 0x800) {
      o += (char)(0xC0 | (c >> 6));
      o += (char)(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      o += (char)(0xE0 | (c >> 12));
      o += (char)(0x80 | ((c >> 6) & 0x3F));
      o += (char)(0x80 | (c & 0x3F));
    } else {
      o += (char)(0xF0 | (c >> 18));
      o += (char)(0x80 | ((c >> 12) & 0x3F));
      o += (char)(0x80 | ((c >> 6) & 0x3F));
      o += (char)(0x80 | (c & 0x3F));
    }
  }
  return o;
}

// parseInt(s, 10) after the "element vertex " prefix: NaN when no digits
double js_parse_int(const U32s &line, size_t from) {
  const U32s s = js_trim(U32s(line.begin() + (long)from, line.end()));
  size_t i = 0;
  bool neg = false;
  if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
  double v = 0.0;
  bool any = false;
  while (i < s.size() && s[i] >= '0' && s[i] <= '9') {
    v = v * 10.0 + (double)(s[i] - '0');
    any = true;
    i++;
  }
  if (!any) return NAN;
  return neg ? -v : v;
}

struct PlyProp {
  std::string type, name;
};

struct PlyHeader {
  double count = 0;
  std::vector<PlyProp> props;
  size_t data_offset = 0;
};

PlyHeader parse_header(const uint8_t *bytes, size_t len) {
  const U32s text = utf8_decode(bytes, std::min(kHeaderMax, len));
  static const char kEnd[] = "end_header\n";
  const size_t m = sizeof(kEnd) - 1;
  size_t end = SIZE_MAX;
  for (size_t i = 0; i + m <= text.size() && end == SIZE_MAX; i++) {
    size_t k = 0;
    while (k < m && text[i + k] == (uint8_t)kEnd[k]) k++;
    if (k == m) end = i;
  }
  if (end == SIZE_MAX) fail(DL_ERR_INVALID, "PLY: end_header not found in first 64KB");
  PlyHeader h;
  // the reference indexes bytes with the UTF-16 index of the decoded text
  size_t u16 = 0;
  for (size_t i = 0; i < end; i++) u16 += text[i] > 0xFFFF ? 2 : 1;
  h.data_offset = u16 + m;
  std::vector<U32s> lines;
  U32s cur;
  for (size_t i = 0; i <= end; i++) {
    if (i == end || text[i] == '\n') {
      U32s t = js_trim(cur);
      if (!t.empty()) lines.push_back(std::move(t));
      cur.clear();
    } else {
      cur.push_back(text[i]);
    }
  }
  if (lines.empty() || to_utf8(lines[0]) != "ply") fail(DL_ERR_INVALID, "PLY: missing magic");
  bool binary = false, little = true, in_vertex = false;
  for (size_t li = 1; li < lines.size(); li++) {
    const U32s &line = lines[li];
    if (starts_with(line, "format ")) {
      binary = contains(line, "binary");
      little = contains(line, "little_endian");
      continue;
    }
    if (starts_with(line, "element vertex ")) {
      h.count = js_parse_int(line, strlen("element vertex "));
      in_vertex = true;
      continue;
    }
    if (starts_with(line, "element ")) {
      in_vertex = false;
      continue;
    }
    if (in_vertex && starts_with(line, "property ")) {
      // line.split(/\s+/): parts[1] = type, parts[2] = name
      std::vector<U32s> parts(1);
      for (size_t i = 0; i < line.size();) {
        if (js_space(line[i])) {
          parts.emplace_back();
          while (i < line.size() && js_space(line[i])) i++;
        } else {
          parts.back().push_back(line[i++]);
        }
      }
      const std::string t = parts.size() > 1 ? to_utf8(parts[1]) : "undefined";
      const std::string name = parts.size() > 2 ? to_utf8(parts[2]) : "undefined";
      if (t != "float" && t != "uchar" && t != "int")
        fail(DL_ERR_INVALID, "PLY: unsupported property type \"" + t + "\"");
      h.props.push_back({t, name});
    }
  }
  if (!binary) fail(DL_ERR_INVALID, "PLY: ASCII PLY not supported (binary only)");
  if (!little) fail(DL_ERR_INVALID, "PLY: only little-endian supported");
  // count <= 0, and a count without digits (NaN in the reference)
  if (!(h.count > 0)) fail(DL_ERR_INVALID, "PLY: no vertices");
  return h;
}

std::string js_number(double v) {
  char b[64];
  snprintf(b, sizeof b, "%.0f", v);
  return b;
}

// ------------------------------------------------------------ de-interleave --
struct PlyCols {
  uint32_t P;      // floats per vertex
  uint32_t count;  // vertices
  uint32_t spb;    // vertices per workgroup
  uint32_t x, y, z, s0, s1, s2, d0, d1, d2, op, r0, r1, r2, r3;
  uint32_t nsh;    // f_rest values copied per vertex (3 * coeffs per channel)
  uint32_t flip;
};

// Math.hypot of four f64 values as the JS engines compute it (ply.ts:220):
// +inf if any is infinite, NaN if any is NaN, 0 if all are 0, else the
// max-scaled, Kahan-compensated sum of squares.
__device__ inline double js_hypot4(double a, double b, double c, double d) {
  a = fabs(a);
  b = fabs(b);
  c = fabs(c);
  d = fabs(d);
  if (isinf(a) || isinf(b) || isinf(c) || isinf(d)) return INFINITY;
  if (isnan(a) || isnan(b) || isnan(c) || isnan(d)) return NAN;
  const double mx = fmax(fmax(a, b), fmax(c, d));
  if (mx == 0.0) return 0.0;
  const double v[4] = {a, b, c, d};
  double sum = 0.0, comp = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double q = v[k] / mx;
    const double summand = q * q - comp;
    const double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return __builtin_sqrt(sum) * mx;
}

// One workgroup per spb vertices: their rows are staged in LDS with
// coalesced loads, then each thread converts one vertex; the f_rest copy is
// strided over the workgroup so its writes stay coalesced.
__global__ __launch_bounds__(kPlyBlock) void k_ply_deinterleave(const float *__restrict__ body, PlyCols c,
                                                                const uint32_t *__restrict__ sh_cols,
                                                                float *__restrict__ pos, float *__restrict__ rot,
                                                                float *__restrict__ ls, float *__restrict__ op,
                                                                float *__restrict__ dc, float *__restrict__ sh) {
  extern __shared__ float s_rows[];
  const size_t i0 = (size_t)blockIdx.x * c.spb;
  const uint32_t nrows = (uint32_t)min((size_t)c.spb, (size_t)c.count - i0);
  const size_t nf = (size_t)nrows * c.P;
  const float *src = body + i0 * c.P;
  for (size_t e = threadIdx.x; e < nf; e += kPlyBlock) s_rows[e] = src[e];
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < nrows; t += kPlyBlock) {
    const float *r = s_rows + (size_t)t * c.P;
    const size_t i = i0 + t;
    pos[3 * i + 0] = r[c.x];
    pos[3 * i + 1] = c.flip ? -r[c.y] : r[c.y];
    pos[3 * i + 2] = r[c.z];
    ls[3 * i + 0] = r[c.s0];
    ls[3 * i + 1] = r[c.s1];
    ls[3 * i + 2] = r[c.s2];
    op[i] = r[c.op];
    dc[3 * i + 0] = (float)(kShC0 * (double)r[c.d0]);
    dc[3 * i + 1] = (float)(kShC0 * (double)r[c.d1]);
    dc[3 * i + 2] = (float)(kShC0 * (double)r[c.d2]);
    const double w = r[c.r0], qx = r[c.r1], qy = r[c.r2], qz = r[c.r3];
    double n = js_hypot4(w, qx, qy, qz);
    if (n == 0.0 || isnan(n)) n = 1.0;  // `|| 1`: 0 and NaN are falsy
    const double f = c.flip ? -1.0 : 1.0;
    rot[4 * i + 0] = (float)(w / n);
    rot[4 * i + 1] = (float)((f * qx) / n);
    rot[4 * i + 2] = (float)(qy / n);
    rot[4 * i + 3] = (float)((f * qz) / n);
  }
  if (c.nsh) {
    const size_t total = (size_t)nrows * c.nsh;
    float *dst = sh + i0 * c.nsh;
    for (size_t e = threadIdx.x; e < total; e += kPlyBlock) {
      const uint32_t t = (uint32_t)(e / c.nsh), k = (uint32_t)(e % c.nsh);
      dst[e] = s_rows[(size_t)t * c.P + sh_cols[k]];
    }
  }
}

// ------------------------------------------------------------------ bounds --
// splatBounds (ply.ts:242-255): per coordinate, a strict `<` (`>`) scan in
// index order from +Infinity (-Infinity). Each step keeps the earlier value
// on ties and skips NaN, so an ordered pairwise reduction (earlier range
// first) gives the scan's result bit for bit, zero signs included.
struct MinMax3 {
  float lo[3], hi[3];
};

__device__ inline void mm_take(MinMax3 &a, const MinMax3 &later) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (later.lo[k] < a.lo[k]) a.lo[k] = later.lo[k];
    if (later.hi[k] > a.hi[k]) a.hi[k] = later.hi[k];
  }
}

__device__ inline MinMax3 mm_empty() {
  MinMax3 m;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    m.lo[k] = INFINITY;
    m.hi[k] = -INFINITY;
  }
  return m;
}

// ordered workgroup reduction: thread t's value covers an earlier range than
// thread t + 1's
__device__ inline MinMax3 mm_block_reduce(MinMax3 v, MinMax3 *s) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t w = 1; w < kPlyBlock; w <<= 1) {
    if ((threadIdx.x % (2 * w)) == 0 && threadIdx.x + w < kPlyBlock) mm_take(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  return s[0];
}

// Stage 1: workgroup b scans the contiguous range [b * per, (b + 1) * per);
// each thread a contiguous sub-range, in index order.
__global__ __launch_bounds__(kPlyBlock) void k_bounds_partial(const float *__restrict__ pos, uint32_t n, uint32_t per,
                                                              MinMax3 *__restrict__ part) {
  __shared__ MinMax3 s[kPlyBlock];
  const size_t b0 = (size_t)blockIdx.x * per;
  const size_t b1 = min((size_t)n, b0 + per);
  const size_t sub = (per + kPlyBlock - 1) / kPlyBlock;
  const size_t t0 = b0 + threadIdx.x * sub, t1 = min(b1, t0 + sub);
  MinMax3 m = mm_empty();
  for (size_t i = t0; i < t1; i++) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float v = pos[3 * i + k];
      if (v < m.lo[k]) m.lo[k] = v;
      if (v > m.hi[k]) m.hi[k] = v;
    }
  }
  m = mm_block_reduce(m, s);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(kPlyBlock) void k_bounds_final(const MinMax3 *__restrict__ part, uint32_t nparts,
                                                            MinMax3 *__restrict__ out) {
  __shared__ MinMax3 s[kPlyBlock];
  const uint32_t sub = (nparts + kPlyBlock - 1) / kPlyBlock;
  const uint32_t t0 = threadIdx.x * sub, t1 = min(nparts, t0 + sub);
  MinMax3 m = mm_empty();
  for (uint32_t i = t0; i < t1; i++) mm_take(m, part[i]);
  m = mm_block_reduce(m, s);
  if (threadIdx.x == 0) *out = m;
}

}  // namespace

void Scene::load_ply(const uint8_t *bytes, size_t len, bool flip_y) {
  if (!bytes) fail(DL_ERR_INVALID, "PLY: null source");
  const PlyHeader h = parse_header(bytes, len);
  std::map<std::string, uint32_t> idx;
  for (size_t p = 0; p < h.props.size(); p++) idx[h.props[p].name] = (uint32_t)p;
  for (const PlyProp &p : h.props)
    if (p.type != "float") fail(DL_ERR_INVALID, "PLY: mixed property types not yet supported");
  const uint32_t P = (uint32_t)h.props.size();
  const double total = h.count * (double)P * 4.0;
  if ((double)h.data_offset + total > (double)len)
    fail(DL_ERR_INVALID, "PLY: data section truncated (need " + js_number(total) + " bytes, have " +
                             js_number((double)len - (double)h.data_offset) + ")");
  auto need = [&](const char *n) -> uint32_t {
    auto it = idx.find(n);
    if (it == idx.end()) fail(DL_ERR_INVALID, std::string("PLY: missing property \"") + n + "\"");
    return it->second;
  };
  PlyCols c{};
  c.x = need("x");
  c.y = need("y");
  c.z = need("z");
  c.s0 = need("scale_0");
  c.s1 = need("scale_1");
  c.s2 = need("scale_2");
  c.d0 = need("f_dc_0");
  c.d1 = need("f_dc_1");
  c.d2 = need("f_dc_2");
  c.op = need("opacity");
  c.r0 = need("rot_0");
  c.r1 = need("rot_1");
  c.r2 = need("rot_2");
  c.r3 = need("rot_3");
  if (h.count > 4294967295.0) fail(DL_ERR_CAPACITY, "PLY: more than 2^32 - 1 vertices");
  uint32_t k_rest = 0;
  while (idx.count("f_rest_" + std::to_string(k_rest))) k_rest++;
  const uint32_t cpc = k_rest / 3;
  const uint32_t bands_ = cpc >= 15 ? 3 : cpc >= 8 ? 2 : cpc >= 3 ? 1 : 0;
  std::vector<uint32_t> sh_cols;
  if (bands_ > 0)
    for (uint32_t k = 0; k < 3 * cpc; k++) sh_cols.push_back(idx["f_rest_" + std::to_string(k)]);
  const uint32_t n = (uint32_t)h.count;
  c.P = P;
  c.count = n;
  c.spb = std::max<uint32_t>(1, std::min<uint32_t>(kPlyBlock, kPlyLdsFloats / std::max<uint32_t>(P, 1)));
  c.nsh = (uint32_t)sh_cols.size();
  c.flip = flip_y ? 1u : 0u;

  ctx->make_current();
  hipStream_t s = ctx->stream;
  DeviceBuffer body, cols;
  const size_t body_bytes = (size_t)n * P * 4;
  body.reserve(std::max<size_t>(body_bytes, 4));
  cols.reserve(std::max<size_t>(sh_cols.size(), 1) * 4);
  pos.reserve((size_t)n * 12);
  rot.reserve((size_t)n * 16);
  ls.reserve((size_t)n * 12);
  op.reserve((size_t)n * 4);
  dc.reserve((size_t)n * 12);
  sh.reserve(std::max<size_t>((size_t)n * c.nsh * 4, 256));
  ctx->staging.to_device(body.get(), bytes + h.data_offset, body_bytes, s);
  if (!sh_cols.empty())
    DL_HIP(hipMemcpyAsync(cols.get(), sh_cols.data(), sh_cols.size() * 4, hipMemcpyHostToDevice, s));
  const uint32_t grid = div_up(n, c.spb);
  k_ply_deinterleave<<<grid, kPlyBlock, (size_t)c.spb * P * 4, s>>>(body.as<float>(), c, cols.as<uint32_t>(),
                                                                   pos.as<float>(), rot.as<float>(), ls.as<float>(),
                                                                   op.as<float>(), dc.as<float>(), sh.as<float>());
  DL_HIP(hipGetLastError());
  DL_HIP(hipStreamSynchronize(s));
  count = n;
  bands = bands_;
  sh_coeffs = bands_ > 0 ? cpc : 0;
  build_color_records();
}

void Scene::bounds(double lo[3], double hi[3]) {
  ctx->make_current();
  hipStream_t s = ctx->stream;
  MinMax3 m;
  for (int k = 0; k < 3; k++) {
    m.lo[k] = INFINITY;
    m.hi[k] = -INFINITY;
  }
  if (count > 0) {
    const uint32_t per = 1u << 14;
    const uint32_t parts = div_up(count, per);
    DeviceBuffer part, out;
    part.reserve((size_t)parts * sizeof(MinMax3));
    out.reserve(sizeof(MinMax3));
    k_bounds_partial<<<parts, kPlyBlock, 0, s>>>(pos.as<float>(), count, per, part.as<MinMax3>());
    DL_HIP(hipGetLastError());
    k_bounds_final<<<1, kPlyBlock, 0, s>>>(part.as<MinMax3>(), parts, out.as<MinMax3>());
    DL_HIP(hipGetLastError());
    DL_HIP(hipMemcpyAsync(&m, out.get(), sizeof(MinMax3), hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
  }
  for (int k = 0; k < 3; k++) {
    lo[k] = m.lo[k];
    hi[k] = m.hi[k];
  }
}

}  // namespace dl
