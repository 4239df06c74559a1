// ckpt.cpp -- nmesh's checkpoint variables file (variables.bin) and its
// CRC-64, host side.
#include "ckpt.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

namespace nmx {

namespace {

constexpr uint64_t kPoly = 0xc96c5795d7870f42ULL;   // ECMA-182, bit-reversed

// byte table of the reflected CRC: b shifted through 8 division steps
const uint64_t *byte_table() {
  static uint64_t t[256];
  static bool done = false;
  if (!done) {
    for (int b = 0; b < 256; ++b) {
      uint64_t r = (uint64_t)b;
      for (int k = 0; k < 8; ++k) r = (r & 1) ? (r >> 1) ^ kPoly : (r >> 1);
      t[b] = r;
    }
    done = true;
  }
  return t;
}

// a * b mod P, reflected (bit 63 is x^0)
uint64_t gf2_mul(uint64_t a, uint64_t b) {
  uint64_t r = 0;
  for (int i = 63; i >= 0; --i) {
    if ((a >> i) & 1) r ^= b;
    b = (b >> 1) ^ ((b & 1) ? kPoly : 0);
  }
  return r;
}

// x^n mod P, reflected
uint64_t gf2_xpow(uint64_t n) {
  uint64_t r = 1ULL << 63, sq = 1ULL << 62;
  while (n) {
    if (n & 1) r = gf2_mul(r, sq);
    sq = gf2_mul(sq, sq);
    n >>= 1;
  }
  return r;
}

// scalarwave1's evolved variables in the reference's variable-index order
// (AddEvoVar order, nmesh_scalarwave1.c:20-22) and their slot in a block
const char *const kVarName[5] = {"scalarwave1_phi", "scalarwave1_pi", "scalarwave1_cx", "scalarwave1_cy",
                                 "scalarwave1_cz"};
const int kVarSlot[5] = {4, 0, 1, 2, 3};

bool read_line(FILE *fp, std::string &s) {
  s.clear();
  int ch;
  while ((ch = std::fgetc(fp)) != EOF) {
    if (ch == '\n') return true;
    s.push_back((char)ch);
  }
  return !s.empty();
}

}  // namespace

uint64_t crc64(uint64_t crc, const void *buf, size_t nbytes) {
  const uint64_t *t = byte_table();
  const uint8_t *d = (const uint8_t *)buf;
  crc = ~crc;
  while (nbytes--) crc = t[(crc ^ *d++) & 0xff] ^ (crc >> 8);
  return ~crc;
}

uint64_t crc64_combine(uint64_t crc1, uint64_t crc2, uint64_t nbytes2) {
  return crc2 ^ gf2_mul(crc1, gf2_xpow(8 * nbytes2));
}

void crc_device_tables(std::vector<uint64_t> &tab, std::vector<uint64_t> &vtab, std::vector<uint64_t> &klane,
                       std::vector<uint64_t> &aw, std::vector<uint64_t> &ypow) {
  const uint64_t *t0 = byte_table();
  tab.assign(8 * 256, 0);
  for (int b = 0; b < 256; ++b) tab[b] = t0[b];
  for (int k = 1; k < 8; ++k)
    for (int b = 0; b < 256; ++b) {
      const uint64_t p = tab[(k - 1) * 256 + b];
      tab[k * 256 + b] = (p >> 8) ^ t0[p & 0xff];
    }
  const uint64_t x4096 = gf2_xpow(4096);
  vtab.assign(8 * 256, 0);
  for (int j = 0; j < 8; ++j)
    for (int b = 0; b < 256; ++b) vtab[j * 256 + b] = gf2_mul((uint64_t)b << (8 * j), x4096);
  klane.assign(9 * 64, 0);
  aw.assign(9, 0);
  for (int n = 2; n <= 10; ++n) {
    const long W = 5L * n * n * n;
    for (int l = 0; l < 64 && l < W; ++l) {
      const long cnt = (W - l + 63) / 64, last = l + 64 * (cnt - 1);
      klane[(n - 2) * 64 + l] = gf2_xpow(64ULL * (uint64_t)(W - 1 - last));
    }
    aw[n - 2] = gf2_mul(~0ULL, gf2_xpow(64ULL * (uint64_t)W));
  }
  ypow.assign(4 * 256, 0);
  for (int j = 0; j < 4; ++j) {
    const uint64_t Y = gf2_xpow(320ULL << (8 * j));
    ypow[j * 256] = 1ULL << 63;
    for (int d = 1; d < 256; ++d) ypow[j * 256 + d] = gf2_mul(ypow[j * 256 + d - 1], Y);
  }
}

std::string elm_name(const Elm &E) {
  std::string s = std::to_string(E.p) + "_";
  for (int i = 0; i < E.l; ++i) s.push_back((char)('0' + E.path[i]));
  return s;
}

std::string write_variables_bin(const Mesh &M, const double *blk, const std::string &path) {
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) return "cannot open " + path + " for writing";
  std::fprintf(fp, "number of variables and variable list in this file:\n");
  std::fprintf(fp, "%d\n", 5);
  for (int k = 0; k < 5; ++k) std::fprintf(fp, "%s\n", kVarName[k]);
  std::fprintf(fp, "\n\n");
  std::fprintf(fp, "variable-list-data:\n\n");
  std::vector<unsigned char> buf;
  for (const Elm &E : M.e) {
    const long np = (long)E.n * E.n * E.n;
    std::fprintf(fp, "{\n%s\n%ld\n", elm_name(E).c_str(), np);
    buf.resize(8 * np);
    for (int k = 0; k < 5; ++k) {
      std::fprintf(fp, "%d\n", k);
      const double *v = blk + 5 * E.off + kVarSlot[k] * np;
      for (long q = 0; q < np; ++q) {   // little endian whatever the host
        uint64_t bits;
        std::memcpy(&bits, v + q, 8);
        for (int i = 0; i < 8; ++i) buf[8 * q + i] = (unsigned char)(bits >> (8 * i));
      }
      std::fwrite(buf.data(), 1, buf.size(), fp);
      std::fprintf(fp, "\n");
    }
    std::fprintf(fp, "}\n\n");
  }
  std::fprintf(fp, "EndOfFile.\n");
  if (std::ferror(fp)) { std::fclose(fp); return "write error on " + path; }
  if (std::fclose(fp) != 0) return "write error on " + path;
  return "";
}

std::string read_variables_bin(const Mesh &M, double *blk, const std::string &path) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) return "cannot open " + path;
  struct Closer { FILE *f; ~Closer() { std::fclose(f); } } closer{fp};
  std::string line;
  // header: comment, number of variables, their names (checkpoint_make_vl)
  if (!read_line(fp, line)) return path + ": empty file";
  int nvars = 0;
  if (!read_line(fp, line) || std::sscanf(line.c_str(), "%d", &nvars) != 1 || nvars < 0 || nvars > 100000)
    return path + ": bad number of variables";
  std::vector<int> slot(nvars, -1);
  for (int i = 0; i < nvars; ++i) {
    if (!read_line(fp, line)) return path + ": truncated variable list";
    for (int k = 0; k < 5; ++k)
      if (line == kVarName[k]) slot[i] = kVarSlot[k];
  }
  while (read_line(fp, line) && line != "variable-list-data:") {}
  std::unordered_map<std::string, int> index;
  for (size_t e = 0; e < M.e.size(); ++e) index[elm_name(M.e[e])] = (int)e;
  std::vector<int> have(M.e.size(), 0);
  std::vector<unsigned char> buf;
  bool end = false;
  while (read_line(fp, line)) {
    if (line.empty()) continue;
    if (line == "EndOfFile.") { end = true; break; }
    if (line != "{") return path + ": expected '{', found '" + line.substr(0, 40) + "'";
    std::string name;
    if (!read_line(fp, name) || !read_line(fp, line)) return path + ": truncated element header";
    const long np = std::atol(line.c_str());
    auto it = index.find(name);
    if (it == index.end()) return path + ": element " + name + " is not in this mesh";
    const Elm &E = M.e[it->second];
    const long mynp = (long)E.n * E.n * E.n;
    if (np != mynp)
      return path + ": element " + name + " has " + std::to_string(np) + " points, the mesh " +
             std::to_string(mynp);
    for (;;) {
      if (!read_line(fp, line)) return path + ": truncated data of element " + name;
      if (line == "}") break;
      int vli = -1, n0 = 0, n1 = 0, n2 = 0;
      const int got = std::sscanf(line.c_str(), "%d %d %d %d", &vli, &n0, &n1, &n2);
      if (got != 1 && got != 4) return path + ": bad variable line in element " + name;
      if (vli < 0 || vli >= nvars) return path + ": variable index out of range in element " + name;
      const long N = (got == 4) ? (long)n0 * n1 * n2 : np;
      if (N < 0 || N > (1L << 31)) return path + ": bad array size in element " + name;
      buf.resize(8 * N);
      if (std::fread(buf.data(), 1, buf.size(), fp) != buf.size())
        return path + ": truncated array in element " + name;
      if (std::fgetc(fp) != '\n') return path + ": missing newline after an array of element " + name;
      const int s = slot[vli];
      if (s < 0) continue;   // a variable this build does not evolve
      if (N != np) return path + ": array of element " + name + " has the wrong size";
      double *v = blk + 5 * E.off + s * np;
      for (long q = 0; q < np; ++q) {
        uint64_t bits = 0;
        for (int i = 0; i < 8; ++i) bits |= (uint64_t)buf[8 * q + i] << (8 * i);
        std::memcpy(v + q, &bits, 8);
      }
      have[it->second] |= 1 << s;
    }
  }
  if (!end) return path + ": no EndOfFile.";
  for (size_t e = 0; e < M.e.size(); ++e)
    if (have[e] != 31) return path + ": element " + elm_name(M.e[e]) + " lacks some of the five variables";
  return "";
}

}  // namespace nmx
