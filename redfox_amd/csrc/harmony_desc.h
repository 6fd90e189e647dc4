// harmony_desc.h - the per-instance description of a Harmony kernel.
//
// Red Fox bakes each kernel instance's tuple types into the source it
// generates (RelationalAlgebraKernel::cudaSourceRepresentation,
// RelationalAlgebraKernel.cpp:849-909: elementType() of the destination /
// source variables, _domains / _keyFields, the PermuteMap, the comparison)
// and stores the compiled result in the HIR kernel's `code` bytes
// (HarmonyIR.proto Kernel.code; RelationalAlgebraCompiler.cpp calls
// compilePTXSource / compileBINSource for it).  The HIR variables
// themselves carry no tuple information: every relation buffer is declared
// I8 with a byte size (RelationalAlgebraCompiler.cpp:1135,1158).
//
// For gfx950 the same members are written as text into `code`:
//
//   rfx1 <family> [d=T] [d1=T] [a=T] [b=T] [c=T] [e=T] [dom=N] [map=i,j,..]
//        [agg=total|min|max|count] [cmp=FIELD:OP:CONST] [parts=N] [nt=N]
//
// where T is a tuple: its fields' bit widths (Element::limit) separated by
// commas, each with an optional kind letter (i integer, the default;
// f float; p pointer = a string address), or "-" for an empty variable.
// d, d1, a, b, c, e are the instance's destination, destination_1,
// sourceA, sourceB, sourceC and sourceD as the RelationalAlgebraKernel
// constructors assign them (RelationalAlgebraKernel.cpp:690-818).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace rfx {

struct TupleDesc {
  int n = 0;
  uint32_t bits[16] = {};
  char kind[16] = {};

  uint32_t total() const {
    uint32_t t = 0;
    for (int i = 0; i < n; ++i) t += bits[i];
    return t;
  }
  // Bits of the fields after field i (Tuple.h BitsAfter).
  uint32_t bits_after(int i) const {
    uint32_t t = 0;
    for (int j = i + 1; j < n; ++j) t += bits[j];
    return t;
  }
  // Storage bytes of the tuple's BasicType: the smallest of 8, 16, 32, 64
  // or 128 bits that holds every field (Tuple.h:30-37); 0 past 128.
  int bytes() const {
    const uint32_t t = total();
    if (t <= 8) return 1;
    if (t <= 16) return 2;
    if (t <= 32) return 4;
    if (t <= 64) return 8;
    if (t <= 128) return 16;
    return 0;
  }
  // The reference's bytes(): (sum of limits + 7) / 8.
  int packed_bytes() const { return static_cast<int>((total() + 7) / 8); }
  bool is_string() const { return n == 1 && kind[0] == 'p'; }
  bool is_f64() const { return n == 1 && kind[0] == 'f' && bits[0] == 64; }
};

struct CmpDesc {
  int field = 0;
  int op = 0;  // RFX_CMP_*
  uint64_t constant = 0;
};

struct KernelDesc {
  std::string family;
  TupleDesc d, d1, a, b, c, e;
  bool has_d = false, has_d1 = false, has_a = false, has_b = false, has_c = false, has_e = false;
  int dom = -1;
  std::vector<uint32_t> map;
  std::string agg;
  bool has_cmp = false;
  CmpDesc cmp;
  int parts = 128;
  int nt = 128;
};

// True when `code` holds a descriptor (starts with "rfx1 ").
inline bool is_desc(const char* code, uint64_t len) {
  return code && len >= 5 && std::memcmp(code, "rfx1 ", 5) == 0;
}

inline bool parse_tuple(const std::string& v, TupleDesc* t) {
  *t = TupleDesc{};
  if (v == "-" || v.empty()) return true;
  size_t i = 0;
  while (i <= v.size()) {
    size_t j = v.find(',', i);
    if (j == std::string::npos) j = v.size();
    std::string f = v.substr(i, j - i);
    if (f.empty() || t->n == 16) return false;
    char k = 'i';
    if (f.back() == 'i' || f.back() == 'f' || f.back() == 'p') {
      k = f.back();
      f.pop_back();
    }
    char* end = nullptr;
    const unsigned long b = std::strtoul(f.c_str(), &end, 10);
    if (f.empty() || *end || b == 0 || b > 128) return false;
    t->bits[t->n] = static_cast<uint32_t>(b);
    t->kind[t->n] = k;
    ++t->n;
    i = j + 1;
  }
  return true;
}

// Comparison names of the select descriptor, in RFX_CMP_* order.
inline int cmp_code(const std::string& s) {
  static const char* names[] = {"eq", "ne", "lt", "le", "gt", "ge", "slt",
                                "sle", "sgt", "sge", "dlt", "dle", "dgt", "dge"};
  for (int i = 0; i < 14; ++i)
    if (s == names[i]) return i;
  return -1;
}

// Parse a descriptor; on failure *err names the offending token.
inline bool parse_desc(const char* code, uint64_t len, KernelDesc* out, std::string* err) {
  *out = KernelDesc{};
  std::string text(code, code + len);
  std::vector<std::string> tok;
  size_t i = 0;
  while (i < text.size()) {
    while (i < text.size() && (text[i] == ' ' || text[i] == '\n' || text[i] == '\t' || text[i] == '\0')) ++i;
    size_t j = i;
    while (j < text.size() && !(text[j] == ' ' || text[j] == '\n' || text[j] == '\t' || text[j] == '\0')) ++j;
    if (j > i) tok.push_back(text.substr(i, j - i));
    i = j;
  }
  if (tok.size() < 2 || tok[0] != "rfx1") {
    *err = "not an rfx1 descriptor";
    return false;
  }
  out->family = tok[1];
  for (size_t t = 2; t < tok.size(); ++t) {
    const size_t eqp = tok[t].find('=');
    if (eqp == std::string::npos) {
      *err = tok[t];
      return false;
    }
    const std::string k = tok[t].substr(0, eqp), v = tok[t].substr(eqp + 1);
    bool ok = true;
    if (k == "d") ok = out->has_d = parse_tuple(v, &out->d);
    else if (k == "d1") ok = out->has_d1 = parse_tuple(v, &out->d1);
    else if (k == "a") ok = out->has_a = parse_tuple(v, &out->a);
    else if (k == "b") ok = out->has_b = parse_tuple(v, &out->b);
    else if (k == "c") ok = out->has_c = parse_tuple(v, &out->c);
    else if (k == "e") ok = out->has_e = parse_tuple(v, &out->e);
    else if (k == "dom") out->dom = std::atoi(v.c_str());
    else if (k == "parts") out->parts = std::atoi(v.c_str());
    else if (k == "nt") out->nt = std::atoi(v.c_str());
    else if (k == "agg") out->agg = v;
    else if (k == "map") {
      size_t p = 0;
      while (p < v.size()) {
        size_t q = v.find(',', p);
        if (q == std::string::npos) q = v.size();
        out->map.push_back(static_cast<uint32_t>(std::strtoul(v.substr(p, q - p).c_str(), nullptr, 10)));
        p = q + 1;
      }
    } else if (k == "cmp") {
      const size_t c1 = v.find(':'), c2 = c1 == std::string::npos ? c1 : v.find(':', c1 + 1);
      if (c2 == std::string::npos) ok = false;
      else {
        out->cmp.field = std::atoi(v.substr(0, c1).c_str());
        out->cmp.op = cmp_code(v.substr(c1 + 1, c2 - c1 - 1));
        out->cmp.constant = std::strtoull(v.substr(c2 + 1).c_str(), nullptr, 0);
        ok = out->cmp.op >= 0;
        out->has_cmp = ok;
      }
    } else {
      ok = false;
    }
    if (!ok) {
      *err = tok[t];
      return false;
    }
  }
  return true;
}

}  // namespace rfx
