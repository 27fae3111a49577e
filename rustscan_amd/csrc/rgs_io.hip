// Scene files: RustGS's splat PLY and .splat formats (RustGS/src/io/scene_io.rs),
// read and written on the host, byte for byte as the reference does.
//
// PLY writer (save_splats_ply, scene_io.rs:485-550, header :99-179): ASCII,
// the 3DGS vertex layout (x y z nx ny nz f_dc_0..2 f_rest_0..44 opacity
// scale_0..2 rot_0..3), raw opacity logits, log scales, f_rest channel-major
// (:216-226) and padded or cut to 45 values, every float in Rust's `{}`
// form (shortest round-trip digits, no exponent).
// PLY reader (load_splats_ply, :629-757, header :241-360, fields :362-400,
// degree :402-432, rows :434-475): ASCII or binary little-endian float
// properties in any order, opacity as raw logit or probability, f_rest
// channel-major or interleaved, SH degree from the comment or from the
// f_rest count, degrees above 3 rejected.
// .splat (save_splats_splat :552-615, load_splats_splat :759-870): 32-byte
// rows of position, scale, RGBA8 and a quantised quaternion.
//
// Large files: the ASCII body is parsed, and written, by up to 16 threads
// over line-aligned chunks; rows keep their file order.
#include <locale.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "rgs_host.h"

struct rgs_scene {
  rgs_splat_metadata meta{};
  int64_t n = 0;
  int sh_degree = 0;
  std::vector<float> positions, log_scales, rotations, opacity_logits, sh_coeffs;
};

namespace {

thread_local std::string g_io_err;

int io_fail(rgs_ctx* c, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c)
    c->err = buf;
  else
    g_io_err = buf;
  return code;
}

constexpr float kShC0 = 0.2820948f;  // sh.rs:21

int sh_coeff_count(int degree) { return (degree + 1) * (degree + 1); }

unsigned io_threads(size_t work_items) {
  unsigned t = std::thread::hardware_concurrency();
  t = std::max(1u, std::min(t, 16u));
  const size_t by_work = std::max<size_t>(1, work_items / 4096);
  return (unsigned)std::min<size_t>(t, by_work);
}

// Runs f(i) for i in [0, parts) on up to `parts` threads.
template <class F>
void parallel_parts(unsigned parts, F&& f) {
  if (parts <= 1) {
    f(0u);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(parts);
  for (unsigned i = 0; i < parts; ++i) th.emplace_back([&f, i] { f(i); });
  for (auto& t : th) t.join();
}

// ---- Rust float text -------------------------------------------------------

// Rust's `{}` of an f32: the shortest digits that read back as the same float,
// positional (no exponent); "NaN", "inf", "-inf"; "-0" for negative zero.
void put_f32(std::string& out, float v) {
  if (std::isnan(v)) {
    out += "NaN";
    return;
  }
  if (std::isinf(v)) {
    out += v < 0 ? "-inf" : "inf";
    return;
  }
  char buf[48];
  const auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
  const char* p = buf;
  if (*p == '-') {
    out += '-';
    ++p;
  }
  const char* e = std::find(p, (const char*)r.ptr, 'e');
  char digits[16];
  int nd = 0;
  for (const char* q = p; q < e; ++q)
    if (*q != '.') digits[nd++] = *q;
  int exp10 = 0;
  std::from_chars(e + 1 + (e[1] == '+' ? 1 : 0), r.ptr, exp10);
  if (nd == 1 && digits[0] == '0') {
    out += '0';
    return;
  }
  const int point = exp10 + 1;  // digits before the decimal point
  if (point <= 0) {
    out += "0.";
    out.append((size_t)(-point), '0');
    out.append(digits, nd);
  } else if (point >= nd) {
    out.append(digits, nd);
    out.append((size_t)(point - nd), '0');
  } else {
    out.append(digits, point);
    out += '.';
    out.append(digits + point, nd - point);
  }
}

locale_t c_locale() {
  static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
  return loc;
}

bool ieq(const char* b, const char* e, const char* word) {
  const size_t n = strlen(word);
  if ((size_t)(e - b) != n) return false;
  for (size_t i = 0; i < n; ++i)
    if ((char)tolower((unsigned char)b[i]) != word[i]) return false;
  return true;
}

// Rust's `str::parse::<f32>`: [+-] then inf | infinity | nan (any case) or
// digits with an optional point and exponent, at least one mantissa digit;
// correctly rounded (overflow to inf, underflow to subnormals or zero).
bool parse_f32(const char* b, const char* e, float* out) {
  const char* p = b;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    ++p;
  }
  if (p == e) return false;
  if (ieq(p, e, "inf") || ieq(p, e, "infinity")) {
    *out = neg ? -INFINITY : INFINITY;
    return true;
  }
  if (ieq(p, e, "nan")) {
    *out = neg ? -NAN : NAN;
    return true;
  }
  const char* q = p;
  int mant = 0;
  while (q < e && *q >= '0' && *q <= '9') ++q, ++mant;
  if (q < e && *q == '.') {
    ++q;
    while (q < e && *q >= '0' && *q <= '9') ++q, ++mant;
  }
  if (mant == 0) return false;
  if (q < e && (*q == 'e' || *q == 'E')) {
    ++q;
    if (q < e && (*q == '+' || *q == '-')) ++q;
    int ed = 0;
    while (q < e && *q >= '0' && *q <= '9') ++q, ++ed;
    if (ed == 0) return false;
  }
  if (q != e) return false;
  char small[64];
  std::string big;
  const size_t len = (size_t)(e - p);
  const char* s;
  if (len < sizeof(small)) {
    memcpy(small, p, len);
    small[len] = 0;
    s = small;
  } else {
    big.assign(p, len);
    s = big.c_str();
  }
  char* end = nullptr;
  const float v = strtof_l(s, &end, c_locale());
  *out = neg ? -v : v;
  return true;
}

bool is_ws(char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f'; }

void trim(const char*& b, const char*& e) {
  while (b < e && is_ws(*b)) ++b;
  while (e > b && is_ws(e[-1])) --e;
}

// Whole UTF-8 validity, as Rust's str::from_utf8 requires of the header
// lines and of an ASCII body.
bool valid_utf8(const unsigned char* p, size_t n) {
  size_t i = 0;
  while (i < n) {
    const unsigned char c = p[i];
    if (c < 0x80) {
      ++i;
      continue;
    }
    int len;
    uint32_t cp;
    if ((c & 0xE0) == 0xC0) {
      len = 2;
      cp = c & 0x1F;
    } else if ((c & 0xF0) == 0xE0) {
      len = 3;
      cp = c & 0x0F;
    } else if ((c & 0xF8) == 0xF0) {
      len = 4;
      cp = c & 0x07;
    } else {
      return false;
    }
    if (i + len > n) return false;
    for (int k = 1; k < len; ++k) {
      if ((p[i + k] & 0xC0) != 0x80) return false;
      cp = (cp << 6) | (p[i + k] & 0x3F);
    }
    if ((len == 2 && cp < 0x80) || (len == 3 && cp < 0x800) || (len == 4 && cp < 0x10000) || cp > 0x10FFFF ||
        (cp >= 0xD800 && cp <= 0xDFFF))
      return false;
    i += len;
  }
  return true;
}

// `s.parse::<usize>()`: optional '+', decimal digits, no overflow.
bool parse_usize(const char* b, const char* e, uint64_t* out) {
  if (b < e && *b == '+') ++b;
  if (b == e) return false;
  uint64_t v = 0;
  for (const char* p = b; p < e; ++p) {
    if (*p < '0' || *p > '9') return false;
    const uint64_t d = (uint64_t)(*p - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

// ---- files ------------------------------------------------------------------

enum class FileFormat { Ply, Splat };

// splat_file_format (scene_io.rs:188-207): by Path::extension, lower case.
int file_format(rgs_ctx* c, const char* path, FileFormat* out) {
  const char* slash = strrchr(path, '/');
  const char* name = slash ? slash + 1 : path;
  const char* dot = strrchr(name, '.');
  if (!dot || dot == name)
    return io_fail(c, RGS_ERR_INVALID, "invalid scene format: missing splat output extension; expected .splat or .ply");
  std::string ext(dot + 1);
  for (char& ch : ext) ch = (char)tolower((unsigned char)ch);
  if (ext == "ply") {
    *out = FileFormat::Ply;
    return RGS_OK;
  }
  if (ext == "splat") {
    *out = FileFormat::Splat;
    return RGS_OK;
  }
  return io_fail(c, RGS_ERR_INVALID,
                 "invalid scene format: unsupported splat output extension '.%s'; expected .splat or .ply", dot + 1);
}

int read_file(rgs_ctx* c, const char* path, std::vector<char>& bytes) {
  FILE* f = fopen(path, "rb");
  if (!f) return io_fail(c, RGS_ERR_INVALID, "failed to read scene %s: %s", path, strerror(errno));
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  bytes.resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = bytes.empty() ? 0 : fread(bytes.data(), 1, bytes.size(), f);
  fclose(f);
  if (got != bytes.size()) return io_fail(c, RGS_ERR_INVALID, "failed to read scene %s: short read", path);
  return RGS_OK;
}

int write_file(rgs_ctx* c, const char* path, const std::vector<std::string>& parts) {
  FILE* f = fopen(path, "wb");
  if (!f) return io_fail(c, RGS_ERR_INVALID, "failed to write scene %s: %s", path, strerror(errno));
  for (const auto& p : parts)
    if (!p.empty() && fwrite(p.data(), 1, p.size(), f) != p.size()) {
      fclose(f);
      return io_fail(c, RGS_ERR_INVALID, "failed to write scene %s: %s", path, strerror(errno));
    }
  if (fclose(f) != 0) return io_fail(c, RGS_ERR_INVALID, "failed to write scene %s: %s", path, strerror(errno));
  return RGS_OK;
}

struct SplatsIn {
  int64_t n;
  int deg;
  const float *pos, *ls, *rot, *op, *sh;
};

// ---- PLY writer ---------------------------------------------------------------

std::string ply_header(const rgs_splat_metadata& m, int64_t n) {
  std::string h = "ply\nformat ascii 1.0\ncomment Generated by RustGS\n";
  h += "comment iterations " + std::to_string((unsigned long long)m.iterations) + "\n";
  h += "comment final_loss ";
  put_f32(h, m.final_loss);
  h += "\ncomment sh_degree " + std::to_string((unsigned long long)m.sh_degree) + "\n";
  h += "comment opacity_representation raw_logit\ncomment sh_coeff_layout channel_major\n";
  h += "element vertex " + std::to_string((long long)n) + "\n";
  for (const char* p : {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"})
    h += std::string("property float ") + p + "\n";
  for (int i = 0; i < 45; ++i) h += "property float f_rest_" + std::to_string(i) + "\n";
  for (const char* p : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"})
    h += std::string("property float ") + p + "\n";
  h += "end_header\n";
  return h;
}

void ply_row(std::string& o, const SplatsIn& S, int64_t i) {
  const int K = sh_coeff_count(S.deg);
  const float* sh = S.sh + (size_t)i * K * 3;
  auto f = [&](float v, char sep) {
    put_f32(o, v);
    o += sep;
  };
  f(S.pos[i * 3], ' ');
  f(S.pos[i * 3 + 1], ' ');
  f(S.pos[i * 3 + 2], ' ');
  o += "0 0 0 ";
  f(sh[0], ' ');
  f(sh[1], ' ');
  f(sh[2], ' ');
  // interleaved_sh_rest_to_channel_major: value (coefficient j, channel ch) at ch * (K-1) + j
  const int rest = K - 1;
  for (int t = 0; t < 45; ++t) {
    float v = 0.0f;
    if (t < rest * 3) {
      const int ch = t / rest, j = t % rest;
      v = sh[3 + j * 3 + ch];
    }
    f(v, ' ');
  }
  f(S.op[i], ' ');
  f(S.ls[i * 3], ' ');
  f(S.ls[i * 3 + 1], ' ');
  f(S.ls[i * 3 + 2], ' ');
  f(S.rot[i * 4], ' ');
  f(S.rot[i * 4 + 1], ' ');
  f(S.rot[i * 4 + 2], ' ');
  f(S.rot[i * 4 + 3], '\n');
}

int save_ply(rgs_ctx* c, const char* path, const SplatsIn& S, const rgs_splat_metadata& m) {
  const unsigned parts = io_threads((size_t)S.n);
  std::vector<std::string> out(parts + 1);
  out[0] = ply_header(m, S.n);
  parallel_parts(parts, [&](unsigned p) {
    const int64_t a = S.n * p / parts, b = S.n * (p + 1) / parts;
    std::string& o = out[p + 1];
    o.reserve((size_t)(b - a) * 200);
    for (int64_t i = a; i < b; ++i) ply_row(o, S, i);
  });
  return write_file(c, path, out);
}

// ---- .splat -----------------------------------------------------------------

// f32::round: half away from zero.
float round_away(float v) { return std::round(v); }

uint8_t normalized_to_u8(float v) {  // scene_io.rs:880-882
  const float x = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  const float r = round_away(x * 255.0f);
  return r == r ? (uint8_t)r : (uint8_t)0;
}

uint8_t quat_to_u8(float v) {  // scene_io.rs:884-889
  const float x = v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v);
  float r = round_away(x * 128.0f + 128.0f);
  r = r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r);
  return r == r ? (uint8_t)r : (uint8_t)0;
}

float u8_to_quat(uint8_t b) { return ((float)b - 128.0f) / 128.0f; }  // scene_io.rs:891-894

void normalized_quaternion(const float q[4], float out[4]) {  // scene_io.rs:896-910
  float len_sq = 0.0f;
  for (int k = 0; k < 4; ++k) len_sq = len_sq + q[k] * q[k];
  if (!std::isfinite(len_sq) || len_sq <= 1e-12f) {
    out[0] = 1.0f;
    out[1] = out[2] = out[3] = 0.0f;
    return;
  }
  const float inv = 1.0f / std::sqrt(len_sq);
  for (int k = 0; k < 4; ++k) out[k] = q[k] * inv;
}

int save_splat(rgs_ctx* c, const char* path, const SplatsIn& S) {
  const int K = sh_coeff_count(S.deg);
  std::vector<std::string> out(1);
  out[0].resize((size_t)S.n * 32);
  char* o = &out[0][0];
  for (int64_t i = 0; i < S.n; ++i) {
    float f[6] = {S.pos[i * 3], S.pos[i * 3 + 1], S.pos[i * 3 + 2], std::exp(S.ls[i * 3]), std::exp(S.ls[i * 3 + 1]),
                  std::exp(S.ls[i * 3 + 2])};
    memcpy(o, f, 24);
    const float* sh = S.sh + (size_t)i * K * 3;
    float op = 1.0f / (1.0f + std::exp(-S.op[i]));  // sigmoid_scalar (core/splats.rs:322-324), clamped
    op = op < 0.0f ? 0.0f : (op > 1.0f ? 1.0f : op);
    uint8_t b[8];
    for (int ch = 0; ch < 3; ++ch) b[ch] = normalized_to_u8(sh[ch] * kShC0 + 0.5f);  // sh0_to_rgb_value
    b[3] = normalized_to_u8(op);
    float q[4];
    normalized_quaternion(S.rot + i * 4, q);
    for (int k = 0; k < 4; ++k) b[4 + k] = quat_to_u8(q[k]);
    memcpy(o + 24, b, 8);
    o += 32;
  }
  return write_file(c, path, out);
}

float opacity_to_logit(float p) {  // scene_io.rs:182-186
  const float x = p < 1e-6f ? 1e-6f : (p > 1.0f - 1e-6f ? 1.0f - 1e-6f : p);
  return std::log(x / (1.0f - x));
}

int load_splat(rgs_ctx* c, const std::vector<char>& bytes, rgs_scene* sc) {
  if (bytes.size() % 32 != 0)
    return io_fail(c, RGS_ERR_INVALID, "invalid scene format: invalid .splat byte length %zu; expected a multiple of 32",
                   bytes.size());
  const int64_t n = (int64_t)(bytes.size() / 32);
  sc->n = n;
  sc->sh_degree = 0;
  sc->positions.resize(n * 3);
  sc->log_scales.resize(n * 3);
  sc->rotations.resize(n * 4);
  sc->opacity_logits.resize(n);
  sc->sh_coeffs.resize(n * 3);
  for (int64_t i = 0; i < n; ++i) {
    const unsigned char* row = (const unsigned char*)bytes.data() + i * 32;
    float f[6];
    memcpy(f, row, 24);
    for (int k = 0; k < 3; ++k) {
      sc->positions[i * 3 + k] = f[k];
      sc->log_scales[i * 3 + k] = std::log(std::fmax(f[3 + k], 1e-8f));  // f32::max keeps the non-NaN
    }
    for (int ch = 0; ch < 3; ++ch) sc->sh_coeffs[i * 3 + ch] = ((float)row[24 + ch] / 255.0f - 0.5f) / kShC0;
    sc->opacity_logits[i] = opacity_to_logit((float)row[27] / 255.0f);
    const float q[4] = {u8_to_quat(row[28]), u8_to_quat(row[29]), u8_to_quat(row[30]), u8_to_quat(row[31])};
    normalized_quaternion(q, &sc->rotations[i * 4]);
  }
  sc->meta = rgs_splat_metadata{0, 0.0f, (uint64_t)n, 0};
  return RGS_OK;
}

// ---- PLY reader ---------------------------------------------------------------

struct PlyHeader {
  bool binary = false;
  uint64_t vertex_count = 0;
  std::vector<std::string> props;
  int opacity_rep = -1;  // -1 unset, 0 probability, 1 raw logit
  int sh_layout = -1;    // -1 unset, 0 interleaved, 1 channel-major
  rgs_splat_metadata meta{};
  bool explicit_degree = false;
};

bool strip(const char* b, const char* e, const char* prefix, const char** rest) {
  const size_t n = strlen(prefix);
  if ((size_t)(e - b) < n || memcmp(b, prefix, n) != 0) return false;
  *rest = b + n;
  return true;
}

std::vector<std::pair<const char*, const char*>> split_ws(const char* b, const char* e) {
  std::vector<std::pair<const char*, const char*>> t;
  const char* p = b;
  while (p < e) {
    while (p < e && is_ws(*p)) ++p;
    if (p == e) break;
    const char* s = p;
    while (p < e && !is_ws(*p)) ++p;
    t.emplace_back(s, p);
  }
  return t;
}

// parse_ply_header (scene_io.rs:241-360)
int parse_ply_header(rgs_ctx* c, const std::vector<char>& bytes, PlyHeader* h, size_t* body) {
  size_t off = 0;
  bool magic = false;
  std::string element;
  bool have_element = false;
  const char* base = bytes.data();
  while (off < bytes.size()) {
    const char* nl = (const char*)memchr(base + off, '\n', bytes.size() - off);
    const size_t end = nl ? (size_t)(nl - base) + 1 : bytes.size();
    if (!valid_utf8((const unsigned char*)base + off, end - off))
      return io_fail(c, RGS_ERR_INVALID, "parse error: invalid utf-8 in PLY header");
    const char* b = base + off;
    const char* e = base + end;
    off = end;
    trim(b, e);
    if (!magic) {
      if (!(e - b == 3 && memcmp(b, "ply", 3) == 0))
        return io_fail(c, RGS_ERR_INVALID, "invalid scene format: missing PLY magic header");
      magic = true;
      continue;
    }
    const char* rest;
    if (strip(b, e, "format ", &rest)) {
      const char* rb = rest;
      const char* re = e;
      trim(rb, re);
      const std::string f(rb, re);
      if (f == "ascii 1.0")
        h->binary = false;
      else if (f == "binary_little_endian 1.0")
        h->binary = true;
      else
        return io_fail(c, RGS_ERR_INVALID, "invalid scene format: unsupported PLY format %s", f.c_str());
      continue;
    }
    auto value = [&](const char* r) {
      const char* rb = r;
      const char* re = e;
      trim(rb, re);
      return std::make_pair(rb, re);
    };
    if (strip(b, e, "comment iterations ", &rest)) {
      auto v = value(rest);
      uint64_t x;
      h->meta.iterations = parse_usize(v.first, v.second, &x) ? x : 0;
    } else if (strip(b, e, "comment final_loss ", &rest)) {
      auto v = value(rest);
      float x;
      h->meta.final_loss = parse_f32(v.first, v.second, &x) ? x : 0.0f;
    } else if (strip(b, e, "comment sh_degree ", &rest)) {
      auto v = value(rest);
      uint64_t x;
      h->meta.sh_degree = parse_usize(v.first, v.second, &x) ? x : 0;
      h->explicit_degree = true;
    } else if (strip(b, e, "comment SH degree: ", &rest)) {
      auto v = value(rest);
      uint64_t x;
      if (parse_usize(v.first, v.second, &x)) h->meta.sh_degree = x;
      h->explicit_degree = true;
    } else if (strip(b, e, "comment opacity_representation ", &rest)) {
      auto v = value(rest);
      const std::string s(v.first, v.second);
      h->opacity_rep = s == "probability" ? 0 : (s == "raw_logit" ? 1 : -1);
    } else if (strip(b, e, "comment sh_coeff_layout ", &rest)) {
      auto v = value(rest);
      const std::string s(v.first, v.second);
      h->sh_layout = s == "interleaved" ? 0 : (s == "channel_major" ? 1 : -1);
    } else if (strip(b, e, "element ", &rest)) {
      auto parts = split_ws(b, e);
      element = parts.size() > 1 ? std::string(parts[1].first, parts[1].second) : std::string();
      have_element = true;
      uint64_t cnt = 0;
      if (parts.size() > 2 && !parse_usize(parts[2].first, parts[2].second, &cnt)) cnt = 0;
      if (element == "vertex") h->vertex_count = cnt;
    } else if (strip(b, e, "property ", &rest) && have_element && element == "vertex") {
      auto parts = split_ws(b, e);
      const std::string line(b, e);
      if (parts.size() > 1 && std::string(parts[1].first, parts[1].second) == "list")
        return io_fail(c, RGS_ERR_INVALID,
                       "invalid scene format: list vertex properties are not supported in splat PLY loader");
      if (parts.size() != 3)
        return io_fail(c, RGS_ERR_INVALID, "invalid scene format: invalid property declaration: %s", line.c_str());
      const std::string type(parts[1].first, parts[1].second);
      if (type != "float" && type != "float32")
        return io_fail(c, RGS_ERR_INVALID, "invalid scene format: unsupported vertex property type %s", type.c_str());
      h->props.emplace_back(parts[2].first, parts[2].second);
    } else if (e - b == 10 && memcmp(b, "end_header", 10) == 0) {
      *body = off;
      return RGS_OK;
    }
  }
  return io_fail(c, RGS_ERR_INVALID, "invalid scene format: missing end_header in splat PLY");
}

struct FieldMap {
  int pos[3], ls[3], rot[4], op, dc[3];
  std::vector<int> rest;
};

// build_vertex_field_map (scene_io.rs:362-400): the first property of each
// name; f_rest_<k> ordered by k.
int field_map(rgs_ctx* c, const std::vector<std::string>& props, FieldMap* fm) {
  auto find = [&](const char* name, int* out) -> int {
    for (size_t i = 0; i < props.size(); ++i)
      if (props[i] == name) {
        *out = (int)i;
        return RGS_OK;
      }
    return io_fail(c, RGS_ERR_INVALID, "invalid scene format: missing required PLY property %s", name);
  };
  int rc = 0;
  const char* pn[3] = {"x", "y", "z"};
  const char* sn[3] = {"scale_0", "scale_1", "scale_2"};
  const char* rn[4] = {"rot_0", "rot_1", "rot_2", "rot_3"};
  const char* dn[3] = {"f_dc_0", "f_dc_1", "f_dc_2"};
  for (int k = 0; k < 3 && !rc; ++k) rc = find(pn[k], &fm->pos[k]);
  for (int k = 0; k < 3 && !rc; ++k) rc = find(sn[k], &fm->ls[k]);
  if (!rc) rc = find("opacity", &fm->op);
  for (int k = 0; k < 4 && !rc; ++k) rc = find(rn[k], &fm->rot[k]);
  for (int k = 0; k < 3 && !rc; ++k) rc = find(dn[k], &fm->dc[k]);
  if (rc) return rc;
  std::vector<std::pair<uint64_t, int>> rest;
  for (size_t i = 0; i < props.size(); ++i) {
    const char* r;
    const std::string& p = props[i];
    uint64_t k;
    if (strip(p.data(), p.data() + p.size(), "f_rest_", &r) && parse_usize(r, p.data() + p.size(), &k))
      rest.emplace_back(k, (int)i);
  }
  std::stable_sort(rest.begin(), rest.end(), [](auto& a, auto& b) { return a.first < b.first; });
  fm->rest.clear();
  for (auto& x : rest) fm->rest.push_back(x.second);
  return RGS_OK;
}

// infer_sh_degree (scene_io.rs:402-432)
int infer_degree(rgs_ctx* c, bool explicit_deg, uint64_t deg, size_t rest, uint64_t* out) {
  if (explicit_deg) {
    *out = deg;
    return RGS_OK;
  }
  if (rest == 0) {
    *out = 0;
    return RGS_OK;
  }
  if (rest % 3 != 0)
    return io_fail(c, RGS_ERR_INVALID, "invalid scene format: invalid f_rest property count %zu; expected a multiple of 3",
                   rest);
  const uint64_t per = rest / 3 + 1;
  uint64_t d = 0;
  while ((d + 1) * (d + 1) < per) ++d;
  if ((d + 1) * (d + 1) != per)
    return io_fail(c, RGS_ERR_INVALID, "invalid scene format: could not infer SH degree from %zu rest coefficients",
                   rest);
  *out = d;
  return RGS_OK;
}

// append_vertex_values (scene_io.rs:434-475) for row i.
void put_row(rgs_scene* sc, int64_t i, const float* v, const FieldMap& fm, int deg, bool probability,
             bool channel_major) {
  for (int k = 0; k < 3; ++k) sc->positions[i * 3 + k] = v[fm.pos[k]];
  for (int k = 0; k < 3; ++k) sc->log_scales[i * 3 + k] = v[fm.ls[k]];
  for (int k = 0; k < 4; ++k) sc->rotations[i * 4 + k] = v[fm.rot[k]];
  sc->opacity_logits[i] = probability ? opacity_to_logit(v[fm.op]) : v[fm.op];
  const int K = sh_coeff_count(deg);
  float* sh = &sc->sh_coeffs[(size_t)i * K * 3];
  for (int k = 0; k < 3; ++k) sh[k] = v[fm.dc[k]];
  const int keep = (K - 1) * 3;
  if (keep > 0) {
    float rest[72];
    for (int t = 0; t < keep; ++t) rest[t] = t < (int)fm.rest.size() ? v[fm.rest[t]] : 0.0f;
    const int per = keep / 3;
    for (int j = 0; j < per; ++j)
      for (int ch = 0; ch < 3; ++ch)
        sh[3 + j * 3 + ch] = channel_major ? rest[ch * per + j] : rest[j * 3 + ch];
  }
}

int load_ply(rgs_ctx* c, const std::vector<char>& bytes, rgs_scene* sc) {
  PlyHeader h;
  size_t body = 0;
  int rc = parse_ply_header(c, bytes, &h, &body);
  if (rc) return rc;
  FieldMap fm;
  rc = field_map(c, h.props, &fm);
  if (rc) return rc;
  uint64_t deg = 0;
  rc = infer_degree(c, h.explicit_degree, h.meta.sh_degree, fm.rest.size(), &deg);
  if (rc) return rc;
  if (deg > 3)
    return io_fail(c, RGS_ERR_INVALID,
                   "invalid scene format: unsupported sh_degree %llu in splat loader; current path supports up to "
                   "degree 3",
                   (unsigned long long)deg);
  const bool probability = h.opacity_rep == 0;  // default raw logit
  const bool channel_major = h.sh_layout != 0;  // default channel-major
  const size_t P = h.props.size();
  const int K = sh_coeff_count((int)deg);
  auto size_rows = [&](int64_t n) {
    sc->n = n;
    sc->positions.assign(n * 3, 0.0f);
    sc->log_scales.assign(n * 3, 0.0f);
    sc->rotations.assign(n * 4, 0.0f);
    sc->opacity_logits.assign(n, 0.0f);
    sc->sh_coeffs.assign((size_t)n * K * 3, 0.0f);
  };
  int64_t parsed = 0;
  if (h.binary) {
    const size_t row_bytes = P * 4;
    if (h.vertex_count > 0 && row_bytes > 0 && h.vertex_count > (SIZE_MAX / row_bytes))
      return io_fail(c, RGS_ERR_INVALID, "invalid scene format: vertex byte size overflow while reading PLY");
    const size_t expected = (size_t)h.vertex_count * row_bytes;
    const size_t have = bytes.size() - body;
    if (have < expected)
      return io_fail(c, RGS_ERR_INVALID,
                     "invalid scene format: binary PLY truncated: expected at least %zu bytes, got %zu", expected, have);
    parsed = (int64_t)h.vertex_count;
    size_rows(parsed);
    const char* src = bytes.data() + body;
    const unsigned parts = io_threads((size_t)parsed);
    parallel_parts(parts, [&](unsigned p) {
      std::vector<float> v(P);
      const int64_t a = parsed * p / parts, b = parsed * (p + 1) / parts;
      for (int64_t i = a; i < b; ++i) {
        if (P) memcpy(v.data(), src + (size_t)i * row_bytes, row_bytes);
        put_row(sc, i, v.data(), fm, (int)deg, probability, channel_major);
      }
    });
  } else {
    const char* b0 = bytes.data() + body;
    const size_t len = bytes.size() - body;
    if (!valid_utf8((const unsigned char*)b0, len))
      return io_fail(c, RGS_ERR_INVALID, "parse error: invalid utf-8 in PLY body");
    // line-aligned chunks; pass 1 counts the non-empty lines of each
    const unsigned parts = io_threads(len / 256);
    std::vector<size_t> cut(parts + 1, 0);
    cut[parts] = len;
    for (unsigned p = 1; p < parts; ++p) {
      size_t at = len * p / parts;
      const char* nl = at < len ? (const char*)memchr(b0 + at, '\n', len - at) : nullptr;
      cut[p] = nl ? (size_t)(nl - b0) + 1 : len;
      if (cut[p] < cut[p - 1]) cut[p] = cut[p - 1];
    }
    auto for_lines = [&](unsigned p, auto&& fn) {
      size_t off = cut[p];
      while (off < cut[p + 1]) {
        const char* nl = (const char*)memchr(b0 + off, '\n', cut[p + 1] - off);
        const size_t end = nl ? (size_t)(nl - b0) : cut[p + 1];
        const char* lb = b0 + off;
        const char* le = b0 + end;
        off = nl ? end + 1 : cut[p + 1];
        trim(lb, le);
        if (lb == le) continue;
        if (!fn(lb, le)) return;
      }
    };
    std::vector<int64_t> count(parts + 1, 0);
    parallel_parts(parts, [&](unsigned p) {
      int64_t k = 0;
      for_lines(p, [&](const char*, const char*) {
        ++k;
        return true;
      });
      count[p + 1] = k;
    });
    for (unsigned p = 0; p < parts; ++p) count[p + 1] += count[p];
    parsed = count[parts];
    size_rows(parsed);
    // pass 2: parse; the first error in file order wins
    std::vector<std::string> err(parts);
    std::vector<int64_t> err_row(parts, -1);
    parallel_parts(parts, [&](unsigned p) {
      std::vector<float> v(P);
      int64_t row = count[p];
      for_lines(p, [&](const char* lb, const char* le) {
        size_t k = 0;
        const char* q = lb;
        while (q < le) {
          while (q < le && is_ws(*q)) ++q;
          if (q == le) break;
          const char* s = q;
          while (q < le && !is_ws(*q)) ++q;
          float x;
          if (!parse_f32(s, q, &x)) {
            err[p] = "parse error: invalid float literal";
            err_row[p] = row;
            return false;
          }
          if (k < P) v[k] = x;
          ++k;
        }
        if (k != P) {
          char buf[160];
          snprintf(buf, sizeof(buf),
                   "invalid scene format: expected %zu values per vertex in splat PLY, got %zu", P, k);
          err[p] = buf;
          err_row[p] = row;
          return false;
        }
        put_row(sc, row, v.data(), fm, (int)deg, probability, channel_major);
        ++row;
        return true;
      });
    });
    for (unsigned p = 0; p < parts; ++p)
      if (err_row[p] >= 0) return io_fail(c, RGS_ERR_INVALID, "%s", err[p].c_str());
  }
  if ((uint64_t)parsed != h.vertex_count)
    return io_fail(c, RGS_ERR_INVALID, "invalid scene format: vertex count mismatch: header %llu, parsed %lld",
                   (unsigned long long)h.vertex_count, (long long)parsed);
  sc->sh_degree = (int)deg;
  sc->meta = h.meta;
  sc->meta.sh_degree = (uint32_t)deg;
  if (sc->meta.gaussian_count == 0) sc->meta.gaussian_count = (uint64_t)parsed;
  return RGS_OK;
}

int check_splats_in(rgs_ctx* c, const SplatsIn& S) {
  if (S.n < 0 || S.deg < 0 || S.deg > 4) return io_fail(c, RGS_ERR_INVALID, "invalid splat count or SH degree");
  if (S.n > 0 && (!S.pos || !S.ls || !S.rot || !S.op || !S.sh))
    return io_fail(c, RGS_ERR_INVALID, "null splat array");
  return RGS_OK;
}

int save_any(rgs_ctx* c, const char* path, const SplatsIn& S, const rgs_splat_metadata* meta) {
  if (!path) return io_fail(c, RGS_ERR_INVALID, "null path");
  int rc = check_splats_in(c, S);
  if (rc) return rc;
  FileFormat f;
  rc = file_format(c, path, &f);
  if (rc) return rc;
  const rgs_splat_metadata m = meta ? *meta : rgs_splat_metadata{0, 0.0f, (uint64_t)S.n, (uint32_t)S.deg};
  return f == FileFormat::Ply ? save_ply(c, path, S, m) : save_splat(c, path, S);
}

}  // namespace

const char* rgs_io_last_error(void) { return g_io_err.c_str(); }

int rgs_save_splats(rgs_ctx* c, const char* path, int64_t n, int sh_degree, const float* positions,
                    const float* log_scales, const float* rotations, const float* opacity_logits,
                    const float* sh_coeffs, const rgs_splat_metadata* meta) {
  return save_any(c, path, SplatsIn{n, sh_degree, positions, log_scales, rotations, opacity_logits, sh_coeffs}, meta);
}

int rgs_scene_load(rgs_ctx* c, const char* path, rgs_scene** out) {
  if (!path || !out) return io_fail(c, RGS_ERR_INVALID, "null argument");
  *out = nullptr;
  FileFormat f;
  int rc = file_format(c, path, &f);
  if (rc) return rc;
  std::vector<char> bytes;
  rc = read_file(c, path, bytes);
  if (rc) return rc;
  rgs_scene* sc = new rgs_scene();
  rc = f == FileFormat::Ply ? load_ply(c, bytes, sc) : load_splat(c, bytes, sc);
  if (rc) {
    delete sc;
    return rc;
  }
  *out = sc;
  return RGS_OK;
}

int rgs_scene_info(const rgs_scene* sc, int64_t* n, int* sh_degree, rgs_splat_metadata* meta) {
  if (!sc) return RGS_ERR_INVALID;
  if (n) *n = sc->n;
  if (sh_degree) *sh_degree = sc->sh_degree;
  if (meta) *meta = sc->meta;
  return RGS_OK;
}

int rgs_scene_get(const rgs_scene* sc, float* positions, float* log_scales, float* rotations,
                  float* opacity_logits, float* sh_coeffs) {
  if (!sc) return RGS_ERR_INVALID;
  auto cp = [](float* dst, const std::vector<float>& v) {
    if (dst && !v.empty()) memcpy(dst, v.data(), v.size() * 4);
  };
  cp(positions, sc->positions);
  cp(log_scales, sc->log_scales);
  cp(rotations, sc->rotations);
  cp(opacity_logits, sc->opacity_logits);
  cp(sh_coeffs, sc->sh_coeffs);
  return RGS_OK;
}

void rgs_scene_free(rgs_scene* sc) { delete sc; }

int rgs_scene_to_splats(rgs_ctx* c, const rgs_scene* sc, rgs_splats** out) {
  if (!c || !sc || !out) return RGS_ERR_INVALID;
  *out = nullptr;
  if (sc->n < 1) return rgs::fail(c, RGS_ERR_INVALID, "rgs_scene_to_splats: the scene has no splats");
  const int64_t n = sc->n;
  std::vector<float> tf((size_t)n * 10);
  for (int64_t i = 0; i < n; ++i) {
    float* t = &tf[(size_t)i * 10];
    for (int k = 0; k < 3; ++k) t[k] = sc->positions[i * 3 + k];
    for (int k = 0; k < 4; ++k) t[3 + k] = sc->rotations[i * 4 + k];
    for (int k = 0; k < 3; ++k) t[7 + k] = sc->log_scales[i * 3 + k];
  }
  rgs_splats* s = nullptr;
  int rc = rgs_splats_create(c, n, sc->sh_degree, &s);
  if (rc) return rc;
  rc = rgs_splats_import_aos(c, s, tf.data(), sc->sh_coeffs.data(), sc->opacity_logits.data());
  if (rc) {
    rgs_splats_destroy(c, s);
    return rc;
  }
  *out = s;
  return RGS_OK;
}

int rgs_splats_save(rgs_ctx* c, const rgs_splats* s, const char* path, const rgs_splat_metadata* meta) {
  if (!c || !s || !path) return RGS_ERR_INVALID;
  const int64_t n = s->n;
  const int K = s->nk;
  std::vector<float> tf((size_t)n * 10), sh((size_t)n * K * 3), op((size_t)n);
  int rc = rgs_splats_export_aos(c, s, tf.data(), sh.data(), op.data());
  if (rc) return rc;
  std::vector<float> pos((size_t)n * 3), ls((size_t)n * 3), rot((size_t)n * 4);
  for (int64_t i = 0; i < n; ++i) {
    const float* t = &tf[(size_t)i * 10];
    for (int k = 0; k < 3; ++k) pos[i * 3 + k] = t[k];
    for (int k = 0; k < 4; ++k) rot[i * 4 + k] = t[3 + k];
    for (int k = 0; k < 3; ++k) ls[i * 3 + k] = t[7 + k];
  }
  const rgs_splat_metadata m = meta ? *meta : rgs_splat_metadata{0, 0.0f, (uint64_t)n, (uint32_t)s->deg};
  return save_any(c, path, SplatsIn{n, s->deg, pos.data(), ls.data(), rot.data(), op.data(), sh.data()}, &m);
}
