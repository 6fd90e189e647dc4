// hir.cpp - Harmony IR programs: a reader for the HIR file format and an
// executor with Harmony's runtime semantics, whose binary kernels run
// through rfx_harmony_call().
//
// File format (RelationalAlgebraCompiler.cpp:9502-9520 writes it,
// Module.cpp:193-333 reads it): a little-endian u64 byte count, then a
// KernelControlFlowGraph message (HarmonyIR.proto), then `testCount`
// self-test records, each a u64 byte count and a Test message
// (RelationalAlgebraCompiler.cpp:9429-9437).  The messages are decoded here
// from the protobuf wire format directly; no protobuf library is used.
//
// Execution follows Runtime.cpp:
// * every block runs in CFG order (_ZEntry, _ZExit, then the file's
//   blocks; Runtime::execute :70-90); block controls are not executed, as
//   there (ControlLaunch asserts, :655-663);
// * input variables are allocated and filled from `data` or `filename`
//   before the first block (_loadVariables :161-282); a declared size is a
//   byte count there, and so it is here;
// * each block's schedule comes from a backward walk with the block's
//   live-out set (computeLiveInOutSets, DataflowInformation.cpp): before a
//   kernel, every non-input variable it writes with mode Out is allocated
//   with its declared size; after it, every non-input non-output variable
//   it reads with mode In for the last time is freed (_scheduleBasicBlock
//   :303-445).  InOut operands are neither;
// * GetSize writes the source's size word into the target (:1054-1121),
//   Resize allocates a not yet allocated variable with bytes(type) x the
//   size variable's value (:1180-1262), UpdateSize sets the size word
//   (:1271-1290), Free releases the memory (:1298-1314);
// * a binary kernel's operand i is passed as the variable's device address
//   when i is in its family's pointer set, else as the first 8 bytes of the
//   variable (BinaryKernelLaunch::execute :672-1038, getVariable :99-118);
// * a compute (PTX) kernel's operands 0-2 are the values of its ctaCount,
//   threadCount and sharedBytes variables and the rest device addresses
//   (PTXKernelLaunch::execute :506-646);
// * both run through rfx_harmony_launch(), which takes what the generated
//   source baked into the instance (tuple widths, key fields, ...) from an
//   rfx1 descriptor in the kernel's `code` bytes (harmony_desc.h).  A binary
//   kernel without a descriptor goes to rfx_harmony_call(), whose families
//   take those types as one operand past the reference's count.
//
// Differences, each where Harmony's behaviour is undefined or missing:
// a freed variable's address is cleared (Harmony keeps the dangling
// pointer); set_intersection uses set_union's pointer set (Harmony has no
// branch for it and passes uninitialised slots); the self-test records,
// which Harmony loads but never runs (HirManager.cpp:76-163, compiled
// out), are run and compared byte for byte.  Variables take their device
// memory from a pool kept for the life of the program.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "harmony_desc.h"
#include "rfx_internal.h"

namespace rfx {
namespace {

// ---------------------------------------------------------------------------
// Protobuf wire format (varint, 64-bit, length-delimited, 32-bit fields).
// ---------------------------------------------------------------------------
struct Wire {
  const uint8_t* p;
  const uint8_t* end;
  bool ok = true;

  bool more() const { return ok && p < end; }
  uint64_t varint() {
    uint64_t v = 0;
    for (int shift = 0; shift < 64; shift += 7) {
      if (p >= end) {
        ok = false;
        return 0;
      }
      const uint8_t b = *p++;
      v |= static_cast<uint64_t>(b & 0x7F) << shift;
      if (!(b & 0x80)) return v;
    }
    ok = false;
    return 0;
  }
  // Length-delimited payload.
  Wire sub() {
    const uint64_t n = varint();
    if (!ok || n > static_cast<uint64_t>(end - p)) {
      ok = false;
      return Wire{p, p, false};
    }
    Wire w{p, p + n};
    p += n;
    return w;
  }
  std::string bytes() {
    Wire w = sub();
    return ok ? std::string(reinterpret_cast<const char*>(w.p), w.end - w.p) : std::string();
  }
  void skip(int wt) {
    switch (wt) {
      case 0: varint(); break;
      case 1: advance(8); break;
      case 2: sub(); break;
      case 5: advance(4); break;
      default: ok = false;  // groups are not used by HarmonyIR.proto
    }
  }
  void advance(uint64_t n) {
    if (n > static_cast<uint64_t>(end - p)) ok = false;
    else p += n;
  }
};

enum KernelType : int {
  kCompute = 1, kBinary = 2, kControl = 3, kBranch = 4, kGetSize = 5, kResize = 6,
  kUpdateSize = 7, kExit = 8
};
enum Mode : int { kIn = 1, kOut = 2, kInOut = 3 };

struct VarDecl {
  uint32_t name = 0;
  int type = 1;
  uint64_t size = 0;
  bool output = false, input = false;
  std::string data, filename;
  uint32_t seen = 0;  // bit per required field
};

struct Operand {
  int mode = kIn;
  uint32_t var = 0;
};

struct Kernel {
  std::string name;
  uint64_t code_bytes = 0;
  std::string desc;  // the code when it is an rfx1 descriptor; other code is not kept
  int type = 0;
  std::vector<Operand> ops;
  std::vector<uint32_t> targets;
};

struct Block {
  uint32_t id = 0;
  std::vector<Kernel> kernels;
  Kernel control;
};

struct TestRec {
  std::string name, program;
  std::vector<VarDecl> inputs, outputs;
  std::vector<uint32_t> features;
};

bool parse_var(Wire w, VarDecl* v) {
  while (w.more()) {
    const uint64_t key = w.varint();
    const int f = static_cast<int>(key >> 3), wt = static_cast<int>(key & 7);
    if (f == 1 && wt == 0) { v->name = static_cast<uint32_t>(w.varint()); v->seen |= 1; }
    else if (f == 2 && wt == 0) { v->type = static_cast<int>(w.varint()); v->seen |= 2; }
    else if (f == 3 && wt == 0) { v->size = w.varint(); v->seen |= 4; }
    else if (f == 4 && wt == 0) { v->output = w.varint() != 0; v->seen |= 8; }
    else if (f == 5 && wt == 0) { v->input = w.varint() != 0; v->seen |= 16; }
    else if (f == 6 && wt == 2) v->data = w.bytes();
    else if (f == 7 && wt == 2) v->filename = w.bytes();
    else w.skip(wt);
  }
  return w.ok;
}

bool parse_operand(Wire w, Operand* o) {
  uint32_t seen = 0;
  while (w.more()) {
    const uint64_t key = w.varint();
    const int f = static_cast<int>(key >> 3), wt = static_cast<int>(key & 7);
    if (f == 1 && wt == 0) { o->mode = static_cast<int>(w.varint()); seen |= 1; }
    else if (f == 2 && wt == 0) { o->var = static_cast<uint32_t>(w.varint()); seen |= 2; }
    else w.skip(wt);
  }
  return w.ok && seen == 3 && o->mode >= kIn && o->mode <= kInOut;
}

// repeated uint32, packed or not.
bool parse_u32s(Wire& w, int wt, std::vector<uint32_t>* out) {
  if (wt == 0) {
    out->push_back(static_cast<uint32_t>(w.varint()));
  } else if (wt == 2) {
    Wire s = w.sub();
    while (s.more()) out->push_back(static_cast<uint32_t>(s.varint()));
    if (!s.ok) return false;
  } else {
    return false;
  }
  return w.ok;
}

bool parse_kernel(Wire w, Kernel* k) {
  uint32_t seen = 0;
  while (w.more()) {
    const uint64_t key = w.varint();
    const int f = static_cast<int>(key >> 3), wt = static_cast<int>(key & 7);
    if (f == 1 && wt == 2) { k->name = w.bytes(); seen |= 1; }
    else if (f == 2 && wt == 2) {
      const Wire c = w.sub();
      k->code_bytes = static_cast<uint64_t>(c.end - c.p);
      if (is_desc(reinterpret_cast<const char*>(c.p), k->code_bytes))
        k->desc.assign(reinterpret_cast<const char*>(c.p), k->code_bytes);
      seen |= 2;
    }
    else if (f == 3 && wt == 0) { k->type = static_cast<int>(w.varint()); seen |= 4; }
    else if (f == 4 && wt == 2) {
      Operand o;
      if (!parse_operand(w.sub(), &o) || !w.ok) return false;
      k->ops.push_back(o);
    } else if (f == 5) {
      if (!parse_u32s(w, wt, &k->targets)) return false;
    } else {
      w.skip(wt);
    }
  }
  return w.ok && (seen & 5) == 5 && k->type >= kCompute && k->type <= kExit;
}

bool parse_block(Wire w, Block* b) {
  uint32_t seen = 0;
  while (w.more()) {
    const uint64_t key = w.varint();
    const int f = static_cast<int>(key >> 3), wt = static_cast<int>(key & 7);
    if (f == 1 && wt == 0) { b->id = static_cast<uint32_t>(w.varint()); seen |= 1; }
    else if (f == 2 && wt == 2) {
      Kernel k;
      if (!parse_kernel(w.sub(), &k) || !w.ok) return false;
      b->kernels.push_back(std::move(k));
    } else if (f == 3 && wt == 2) {
      if (!parse_kernel(w.sub(), &b->control) || !w.ok) return false;
      seen |= 2;
    } else {
      w.skip(wt);
    }
  }
  return w.ok && seen == 3;
}

bool parse_test(Wire w, TestRec* t) {
  while (w.more()) {
    const uint64_t key = w.varint();
    const int f = static_cast<int>(key >> 3), wt = static_cast<int>(key & 7);
    if ((f == 1 || f == 2) && wt == 2) {
      VarDecl v;
      if (!parse_var(w.sub(), &v) || !w.ok) return false;
      (f == 1 ? t->inputs : t->outputs).push_back(std::move(v));
    } else if (f == 3 && wt == 2) {
      t->program = w.bytes();
    } else if (f == 4 && wt == 2) {
      t->name = w.bytes();
    } else if (f == 5) {
      if (!parse_u32s(w, wt, &t->features)) return false;
    } else {
      w.skip(wt);
    }
  }
  return w.ok;
}

uint32_t type_bytes(int t) {  // VariableDeclaration::bytes
  switch (t) {
    case 1: return 1;
    case 2: return 2;
    case 3: return 4;
    case 4: return 8;
    case 5: return 16;
    case 6: return 4;
    case 7: return 8;
  }
  return 1;
}

// Operand indices a family receives as device addresses
// (BinaryKernelLaunch::execute); the rest are values.
struct Family {
  const char* prefix;
  uint32_t pointers;  // bit i: operand i is an address
};
const Family kFamilies[] = {
    {"set_b40csortpair", 0x3},       {"set_mgpusortpair", 0x3},
    {"set_b40csortkey", 0x1},        {"set_mgpusortkey", 0x1},
    {"set_sort", 0x1},               {"mgpu_join_find_bounds", 0x2F},
    {"mgpu_join_main", 0x1B},        {"mgpu_join_gather", 0xFF},
    {"join_get_result_size", 0x3},   {"set_union", 0x17},
    {"set_difference", 0x17},        {"set_intersection", 0x17},
    {"set_single_reduce", 0x3},      {"set_unique", 0x3},
    {"set_reduce", 0x3F},            {"set_count", 0x3F},
};

const Family* family_of(const std::string& name) {
  for (const Family& f : kFamilies)
    if (name.compare(0, std::strlen(f.prefix), f.prefix) == 0) return &f;
  return nullptr;
}

// Device memory of the variables, pooled by power-of-two capacity (256
// bytes and up) for the life of the program: a variable freed by liveness
// returns its block to the pool and the next allocation of that class
// reuses it.  Every kernel finishes before the next operation runs, so a
// reused block is never still in use.
struct Pool {
  std::map<uint64_t, std::vector<void*>> free;
  std::vector<void*> all;
  static uint64_t cls(uint64_t bytes) {
    uint64_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  int take(uint64_t bytes, void** p, uint64_t* cap) {
    const uint64_t c = cls(bytes);
    auto& v = free[c];
    if (!v.empty()) {
      *p = v.back();
      v.pop_back();
    } else {
      if (hipMalloc(p, c) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(RFX_ENOMEM, "hir: allocating %llu bytes", static_cast<unsigned long long>(c));
      }
      all.push_back(*p);
    }
    *cap = c;
    return RFX_OK;
  }
  void give(void* p, uint64_t cap) { free[cap].push_back(p); }
  void release() {
    for (void* p : all) (void)hipFree(p);
    all.clear();
    free.clear();
  }
};

struct VarState {
  void* addr = nullptr;
  uint64_t block = 0;      // pool class of addr
  uint64_t size_word = 0;  // Harmony's sizeInElements
  uint64_t bytes = 0;      // current logical size in bytes
  uint64_t capacity = 0;   // bytes allocated
};

enum OpKind { kAlloc, kFree, kRun };
struct Op {
  OpKind kind;
  uint32_t var;          // kAlloc / kFree
  const Kernel* kernel;  // kRun
};

}  // namespace
}  // namespace rfx

struct rfx_hir {
  std::string name;
  uint32_t entry = 0, exit_id = 0, test_count = 0;
  std::map<uint32_t, rfx::VarDecl> vars;
  std::vector<rfx::Block> blocks;  // file order
  std::vector<rfx::TestRec> tests;
  // Successors by position in `blocks`; -1 stands for the exit pseudo-block.
  std::vector<std::vector<int>> succ;
  std::vector<std::set<uint32_t>> live_out;
  std::vector<std::vector<rfx::Op>> schedule;
  std::map<uint32_t, rfx::VarState> state;
  std::map<uint32_t, std::string> overrides;  // input contents set by the caller or a test
  rfx::Pool pool;
};

namespace rfx {
namespace {

int decode(const uint8_t* bytes, uint64_t n, rfx_hir* h) {
  if (n < 8) return fail(RFX_EINVAL, "hir: %llu bytes, shorter than the 8-byte prefix",
                         static_cast<unsigned long long>(n));
  uint64_t len = 0;
  std::memcpy(&len, bytes, 8);
  if (len > n - 8) return fail(RFX_EINVAL, "hir: program of %llu bytes in a %llu-byte file",
                               static_cast<unsigned long long>(len),
                               static_cast<unsigned long long>(n));
  Wire w{bytes + 8, bytes + 8 + len};
  uint32_t seen = 0;
  while (w.more()) {
    const uint64_t key = w.varint();
    const int f = static_cast<int>(key >> 3), wt = static_cast<int>(key & 7);
    if (f == 1 && wt == 2) { h->name = w.bytes(); seen |= 1; }
    else if (f == 2 && wt == 2) {
      Block b;
      if (!parse_block(w.sub(), &b) || !w.ok) return fail(RFX_EINVAL, "hir: malformed block");
      h->blocks.push_back(std::move(b));
    } else if (f == 3 && wt == 2) {
      VarDecl v;
      if (!parse_var(w.sub(), &v) || !w.ok || (v.seen & 31) != 31)
        return fail(RFX_EINVAL, "hir: malformed variable");
      if (v.type < 1 || v.type > 7) return fail(RFX_EINVAL, "hir: variable %u has type %d", v.name, v.type);
      if (!h->vars.emplace(v.name, v).second)
        return fail(RFX_EINVAL, "hir: variable %u declared twice", v.name);
    } else if (f == 4 && wt == 0) { h->entry = static_cast<uint32_t>(w.varint()); seen |= 2; }
    else if (f == 5 && wt == 0) { h->exit_id = static_cast<uint32_t>(w.varint()); seen |= 4; }
    else if (f == 6 && wt == 0) { h->test_count = static_cast<uint32_t>(w.varint()); seen |= 8; }
    else w.skip(wt);
  }
  if (!w.ok) return fail(RFX_EINVAL, "hir: malformed program message");
  if (seen != 15) return fail(RFX_EINVAL, "hir: program lacks a required field");
  // Self-test records follow the program.
  uint64_t off = 8 + len;
  for (uint32_t t = 0; t < h->test_count; ++t) {
    if (n - off < 8) return fail(RFX_EINVAL, "hir: test %u missing", t);
    uint64_t tl = 0;
    std::memcpy(&tl, bytes + off, 8);
    off += 8;
    if (tl > n - off) return fail(RFX_EINVAL, "hir: test %u truncated", t);
    TestRec rec;
    if (!parse_test(Wire{bytes + off, bytes + off + tl}, &rec))
      return fail(RFX_EINVAL, "hir: malformed test %u", t);
    h->tests.push_back(std::move(rec));
    off += tl;
  }
  // Resolve ids and operands.
  std::map<uint32_t, int> pos;
  for (size_t i = 0; i < h->blocks.size(); ++i)
    if (!pos.emplace(h->blocks[i].id, static_cast<int>(i)).second)
      return fail(RFX_EINVAL, "hir: block %u declared twice", h->blocks[i].id);
  if (!pos.count(h->entry)) return fail(RFX_EINVAL, "hir: entry block %u not found", h->entry);
  if (!pos.count(h->exit_id)) return fail(RFX_EINVAL, "hir: exit block %u not found", h->exit_id);
  h->succ.assign(h->blocks.size(), {});
  for (size_t i = 0; i < h->blocks.size(); ++i) {
    const Block& b = h->blocks[i];
    auto check_ops = [&](const Kernel& k) {
      for (const Operand& o : k.ops)
        if (!h->vars.count(o.var))
          return fail(RFX_EINVAL, "hir: kernel '%s' names undeclared variable %u", k.name.c_str(), o.var);
      return RFX_OK;
    };
    for (const Kernel& k : b.kernels) {
      RFX_TRY(check_ops(k));
      if ((k.type == kGetSize || k.type == kResize || k.type == kUpdateSize) && k.ops.size() < 2)
        return fail(RFX_EINVAL, "hir: size kernel '%s' needs two operands", k.name.c_str());
    }
    RFX_TRY(check_ops(b.control));
    for (uint32_t t : b.control.targets) {
      auto it = pos.find(t);
      if (it == pos.end()) return fail(RFX_EINVAL, "hir: block %u targets unknown block %u", b.id, t);
      h->succ[i].push_back(it->second);
    }
    if (b.id == h->exit_id) h->succ[i].push_back(-1);
  }
  for (const TestRec& t : h->tests) {
    for (const VarDecl& v : t.inputs)
      if (!h->vars.count(v.name)) return fail(RFX_EINVAL, "hir: test '%s' sets unknown variable %u", t.name.c_str(), v.name);
    for (const VarDecl& v : t.outputs)
      if (!h->vars.count(v.name)) return fail(RFX_EINVAL, "hir: test '%s' checks unknown variable %u", t.name.c_str(), v.name);
  }
  return RFX_OK;
}

// Live-in/live-out sets by iteration to a fixed point
// (DataflowInformation.cpp: defs are Out/InOut operands, uses In/InOut).
void liveness(rfx_hir* h) {
  const size_t nb = h->blocks.size();
  std::vector<std::set<uint32_t>> in(nb);
  h->live_out.assign(nb, {});
  bool changed = true;
  while (changed) {
    changed = false;
    for (size_t i = nb; i-- > 0;) {
      std::set<uint32_t> out;
      for (int s : h->succ[i])
        if (s >= 0) out.insert(in[s].begin(), in[s].end());
      std::set<uint32_t> li = out;
      const auto& ks = h->blocks[i].kernels;
      for (auto k = ks.rbegin(); k != ks.rend(); ++k) {
        for (const Operand& o : k->ops)
          if (o.mode != kIn) li.erase(o.var);
        for (const Operand& o : k->ops)
          if (o.mode != kOut) li.insert(o.var);
      }
      if (li != in[i]) {
        in[i] = std::move(li);
        changed = true;
      }
      h->live_out[i] = std::move(out);
    }
  }
}

// Per-block operation lists (Runtime::_scheduleBasicBlock).
void schedule(rfx_hir* h) {
  h->schedule.assign(h->blocks.size(), {});
  for (size_t i = 0; i < h->blocks.size(); ++i) {
    std::set<uint32_t> live = h->live_out[i];
    std::vector<Op> rev;
    const auto& ks = h->blocks[i].kernels;
    for (auto k = ks.rbegin(); k != ks.rend(); ++k) {
      std::set<uint32_t> allocs, frees;
      for (const Operand& o : k->ops) {
        if (o.mode == kOut && !h->vars[o.var].input) {
          live.erase(o.var);
          if (k->type != kResize && k->type != kUpdateSize) allocs.insert(o.var);
        }
      }
      for (const Operand& o : k->ops) {
        if (o.mode == kIn && !h->vars[o.var].output) {
          if (live.insert(o.var).second && !h->vars[o.var].input) frees.insert(o.var);
        }
      }
      for (uint32_t v : frees) rev.push_back(Op{kFree, v, nullptr});
      rev.push_back(Op{kRun, 0, &*k});
      for (uint32_t v : allocs) rev.push_back(Op{kAlloc, v, nullptr});
    }
    h->schedule[i].assign(rev.rbegin(), rev.rend());
  }
}

// Back to the pool (between runs) or to the device (at destroy).
void release_all(rfx_hir* h, bool destroy = false) {
  for (auto& kv : h->state)
    if (kv.second.addr) h->pool.give(kv.second.addr, kv.second.block);
  h->state.clear();
  if (destroy) h->pool.release();
}

void free_var(rfx_hir* h, uint32_t name) {
  VarState& s = h->state[name];
  if (s.addr) h->pool.give(s.addr, s.block);
  s = VarState{};
}

int alloc_var(rfx_hir* h, uint32_t name, uint64_t bytes) {
  VarState& s = h->state[name];
  if (s.addr) return fail(RFX_EINVAL, "hir: variable %u allocated twice", name);
  void* p = nullptr;
  uint64_t cls = 0;
  RFX_TRY(h->pool.take(bytes ? bytes : 1, &p, &cls));
  s.addr = p;
  s.block = cls;
  s.capacity = bytes;
  return RFX_OK;
}

// First 8 bytes of a variable (Runtime::getVariable).
int value_of(rfx_hir* h, uint32_t name, uint64_t* v) {
  const VarState& s = h->state[name];
  if (!s.addr) return fail(RFX_EINVAL, "hir: reading unallocated variable %u", name);
  *v = 0;
  RFX_HIP(hipMemcpy(v, s.addr, std::min<uint64_t>(8, s.capacity), hipMemcpyDeviceToHost));
  return RFX_OK;
}

int address_of(rfx_hir* h, const Kernel& k, size_t i, uint64_t* slot) {
  const uint32_t v = k.ops[i].var;
  const VarState& s = h->state[v];
  if (!s.addr)
    return fail(RFX_EINVAL, "hir: kernel '%s' operand %zu: variable %u is not allocated",
                k.name.c_str(), i, v);
  *slot = reinterpret_cast<uint64_t>(s.addr);
  return RFX_OK;
}

int load_inputs(rfx_hir* h) {
  for (auto& kv : h->vars) {
    const VarDecl& d = kv.second;
    h->state[d.name] = VarState{};
    if (!d.input) continue;
    std::string buf;
    auto ov = h->overrides.find(d.name);
    if (ov != h->overrides.end()) {
      buf = ov->second;
    } else if (!d.filename.empty()) {
      FILE* f = std::fopen(d.filename.c_str(), "rb");
      if (!f) return fail(RFX_EINVAL, "hir: cannot open '%s' for variable %u", d.filename.c_str(), d.name);
      buf.assign(d.size, '\0');
      const size_t got = d.size ? std::fread(&buf[0], 1, d.size, f) : 0;
      std::fclose(f);
      if (got != d.size) return fail(RFX_EINVAL, "hir: '%s' holds fewer than %llu bytes", d.filename.c_str(),
                                     static_cast<unsigned long long>(d.size));
    } else {
      buf.assign(d.size, '\0');
      std::memcpy(&buf[0], d.data.data(), std::min<size_t>(d.size, d.data.size()));
    }
    RFX_TRY(alloc_var(h, d.name, buf.size()));
    VarState& s = h->state[d.name];
    s.size_word = s.bytes = buf.size();
    if (!buf.empty()) RFX_HIP(hipMemcpy(s.addr, buf.data(), buf.size(), hipMemcpyHostToDevice));
  }
  return RFX_OK;
}

int run_kernel(rfx_hir* h, const Kernel& k) {
  switch (k.type) {
    case kBinary: {
      const Family* fam = family_of(k.name);
      if (!fam) return fail(RFX_ENOTSUP, "hir: no kernel family for binary kernel '%s'", k.name.c_str());
      std::vector<uint64_t> slots(k.ops.size());
      for (size_t i = 0; i < k.ops.size(); ++i) {
        if (i < 32 && ((fam->pointers >> i) & 1)) RFX_TRY(address_of(h, k, i, &slots[i]));
        else RFX_TRY(value_of(h, k.ops[i].var, &slots[i]));
      }
      // With a descriptor the operands are exactly the reference's; without
      // one the compatibility form takes the instance's types as a trailing
      // operand.
      if (!k.desc.empty())
        return rfx_harmony_launch(k.name.c_str(), kBinary, k.desc.data(), k.desc.size(),
                                  slots.data(), static_cast<uint32_t>(slots.size()));
      return rfx_harmony_call(k.name.c_str(), slots.data(), static_cast<uint32_t>(slots.size()));
    }
    case kCompute: {
      // Launch dimensions by value, every other operand by address
      // (PTXKernelLaunch::execute).
      std::vector<uint64_t> slots(k.ops.size());
      for (size_t i = 0; i < k.ops.size(); ++i) {
        if (i < 3) RFX_TRY(value_of(h, k.ops[i].var, &slots[i]));
        else RFX_TRY(address_of(h, k, i, &slots[i]));
      }
      return rfx_harmony_launch(k.name.c_str(), kCompute, k.desc.data(), k.desc.size(), slots.data(),
                                static_cast<uint32_t>(slots.size()));
    }
    case kGetSize: {
      VarState& t = h->state[k.ops[0].var];
      const VarState& src = h->state[k.ops[1].var];
      if (!t.addr || !src.addr)
        return fail(RFX_EINVAL, "hir: get_size '%s' on an unallocated variable", k.name.c_str());
      const uint64_t word = src.size_word;
      RFX_HIP(hipMemcpy(t.addr, &word, std::min<uint64_t>(8, t.capacity), hipMemcpyHostToDevice));
      return RFX_OK;
    }
    case kResize: {
      const uint32_t v = k.ops[0].var;
      uint64_t elems = 0;
      RFX_TRY(value_of(h, k.ops[1].var, &elems));
      if (h->state[v].addr) return fail(RFX_EINVAL, "hir: resize of allocated variable %u", v);
      const uint64_t bytes = elems * type_bytes(h->vars[v].type);
      RFX_TRY(alloc_var(h, v, bytes));
      VarState& s = h->state[v];
      s.size_word = elems;
      s.bytes = bytes;
      return RFX_OK;
    }
    case kUpdateSize: {
      uint64_t value = 0;
      RFX_TRY(value_of(h, k.ops[1].var, &value));
      VarState& s = h->state[k.ops[0].var];
      s.size_word = value;
      s.bytes = std::min(value, s.capacity);
      return RFX_OK;
    }
    default:
      return fail(RFX_ENOTSUP, "hir: kernel '%s' of type %d inside a block", k.name.c_str(), k.type);
  }
}

int run_program(rfx_hir* h) {
  release_all(h);
  RFX_TRY(load_inputs(h));
  for (size_t i = 0; i < h->blocks.size(); ++i) {
    for (const Op& op : h->schedule[i]) {
      if (op.kind == kAlloc) {
        const uint64_t size = h->vars[op.var].size;
        RFX_TRY(alloc_var(h, op.var, size));
        VarState& s = h->state[op.var];
        s.size_word = s.bytes = size;
      } else if (op.kind == kFree) {
        free_var(h, op.var);
      } else {
        RFX_TRY(run_kernel(h, *op.kernel));
      }
    }
  }
  RFX_HIP(hipDeviceSynchronize());
  return RFX_OK;
}

void append(std::string* s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  *s += buf;
}

}  // namespace
}  // namespace rfx

using namespace rfx;

extern "C" {

int rfx_hir_load(rfx_hir** out, const void* bytes, uint64_t nbytes) {
  if (!out || (nbytes && !bytes)) return fail(RFX_EINVAL, "rfx_hir_load: null argument");
  *out = nullptr;
  rfx_hir* h = new rfx_hir;
  const int st = decode(static_cast<const uint8_t*>(bytes), nbytes, h);
  if (st != RFX_OK) {
    delete h;
    return st;
  }
  liveness(h);
  schedule(h);
  *out = h;
  return RFX_OK;
}

void rfx_hir_destroy(rfx_hir* h) {
  if (!h) return;
  (void)hipDeviceSynchronize();
  release_all(h, true);
  delete h;
}

int rfx_hir_info(const rfx_hir* h, rfx_hir_summary* s) {
  if (!h || !s) return fail(RFX_EINVAL, "rfx_hir_info: null argument");
  std::memset(s, 0, sizeof *s);
  std::snprintf(s->name, sizeof s->name, "%s", h->name.c_str());
  s->blocks = static_cast<uint32_t>(h->blocks.size());
  s->variables = static_cast<uint32_t>(h->vars.size());
  for (const Block& b : h->blocks) s->kernels += static_cast<uint32_t>(b.kernels.size());
  s->tests = static_cast<uint32_t>(h->tests.size());
  s->entry = h->entry;
  s->exit = h->exit_id;
  return RFX_OK;
}

int rfx_hir_schedule(const rfx_hir* h, uint32_t block, char* text, uint64_t cap) {
  if (!h || !text || !cap) return fail(RFX_EINVAL, "rfx_hir_schedule: null argument");
  if (block >= h->blocks.size()) return fail(RFX_EINVAL, "rfx_hir_schedule: block %u of %zu", block, h->blocks.size());
  std::string s;
  for (const Op& op : h->schedule[block]) {
    if (!s.empty()) s += "; ";
    if (op.kind == kAlloc) append(&s, "alloc %u", op.var);
    else if (op.kind == kFree) append(&s, "free %u", op.var);
    else append(&s, "run %s", op.kernel->name.c_str());
  }
  if (s.size() + 1 > cap) return fail(RFX_ERANGE, "rfx_hir_schedule: %zu bytes needed", s.size() + 1);
  std::memcpy(text, s.c_str(), s.size() + 1);
  return RFX_OK;
}

int rfx_hir_set_input(rfx_hir* h, uint32_t variable, const void* data, uint64_t bytes) {
  if (!h || (bytes && !data)) return fail(RFX_EINVAL, "rfx_hir_set_input: null argument");
  auto it = h->vars.find(variable);
  if (it == h->vars.end()) return fail(RFX_EINVAL, "rfx_hir_set_input: unknown variable %u", variable);
  if (!it->second.input) return fail(RFX_EINVAL, "rfx_hir_set_input: variable %u is not an input", variable);
  h->overrides[variable].assign(static_cast<const char*>(data), bytes);
  return RFX_OK;
}

int rfx_hir_run(rfx_hir* h) {
  if (!h) return fail(RFX_EINVAL, "rfx_hir_run: null program");
  return run_program(h);
}

int rfx_hir_variable(const rfx_hir* h, uint32_t variable, void** addr, uint64_t* bytes) {
  if (!h || !addr || !bytes) return fail(RFX_EINVAL, "rfx_hir_variable: null argument");
  auto it = h->state.find(variable);
  if (it == h->state.end() || !it->second.addr)
    return fail(RFX_EINVAL, "rfx_hir_variable: variable %u holds no memory", variable);
  *addr = it->second.addr;
  *bytes = it->second.bytes;
  return RFX_OK;
}

int rfx_hir_read(const rfx_hir* h, uint32_t variable, void* dst, uint64_t cap, uint64_t* bytes) {
  void* addr = nullptr;
  uint64_t n = 0;
  RFX_TRY(rfx_hir_variable(h, variable, &addr, &n));
  if (bytes) *bytes = n;
  if (n > cap) return fail(RFX_ERANGE, "rfx_hir_read: variable %u holds %llu bytes", variable,
                           static_cast<unsigned long long>(n));
  if (n) RFX_HIP(hipMemcpy(dst, addr, n, hipMemcpyDeviceToHost));
  return RFX_OK;
}

int rfx_hir_run_tests(rfx_hir* h, uint32_t* passed, uint32_t* failed, char* report, uint64_t cap) {
  if (!h || !passed || !failed) return fail(RFX_EINVAL, "rfx_hir_run_tests: null argument");
  *passed = *failed = 0;
  std::string msg;
  const auto saved = h->overrides;
  for (const TestRec& t : h->tests) {
    h->overrides = saved;
    for (const VarDecl& v : t.inputs) h->overrides[v.name] = v.data;
    const int st = run_program(h);
    bool ok = st == RFX_OK;
    if (!ok) append(&msg, "test '%s': program failed: %s\n", t.name.c_str(), rfx_last_error());
    for (size_t o = 0; ok && o < t.outputs.size(); ++o) {
      const VarDecl& want = t.outputs[o];
      const VarState& s = h->state[want.name];
      if (!s.addr) {
        append(&msg, "test '%s': variable %u holds no memory\n", t.name.c_str(), want.name);
        ok = false;
        break;
      }
      if (s.bytes != want.data.size()) {
        append(&msg, "test '%s': variable %u holds %llu bytes, reference %zu\n", t.name.c_str(),
               want.name, static_cast<unsigned long long>(s.bytes), want.data.size());
        ok = false;
        break;
      }
      std::string got(s.bytes, '\0');
      if (s.bytes && hipMemcpy(&got[0], s.addr, s.bytes, hipMemcpyDeviceToHost) != hipSuccess) {
        append(&msg, "test '%s': reading variable %u failed\n", t.name.c_str(), want.name);
        ok = false;
        break;
      }
      if (got != want.data) {
        size_t at = 0;
        while (at < got.size() && got[at] == want.data[at]) ++at;
        append(&msg, "test '%s': variable %u differs from the reference at byte %zu\n",
               t.name.c_str(), want.name, at);
        ok = false;
      }
    }
    (ok ? *passed : *failed) += 1;
  }
  h->overrides = saved;
  if (report && cap) {
    const size_t n = std::min<size_t>(msg.size(), cap - 1);
    std::memcpy(report, msg.data(), n);
    report[n] = '\0';
  }
  return RFX_OK;
}

}  // extern "C"
