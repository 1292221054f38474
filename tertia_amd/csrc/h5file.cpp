// h5file.cpp — writer and reader of the HDF5 subset described in h5file.h,
// and the C-ABI record helpers qss_h5_write / qss_h5_info / qss_h5_read /
// qss_h5_list (the roles of myhdfshell.cpp's WriteHDF5Record* and
// ReadHDF5Record*).
//
// Structures written (all integers little-endian, offsets and lengths 8 bytes):
//   superblock v0 (96 B) | root object header | local heap (names) |
//   B-tree node (type 0, level 0, one child per symbol-table node) |
//   symbol-table nodes (up to 8 names each, sorted) | dataset object headers
//   (dataspace v1, datatype v1, fill value v2, layout v3 contiguous) | raw data
// Group parameters are the library defaults (leaf K = 4, internal K = 16), so
// nodes are allocated at the sizes a reader derives from the superblock.

#include "h5file.h"

#include <sys/stat.h>

#include <algorithm>
#include <cstring>

#include "qss_internal.h"

namespace h5 {
namespace {

constexpr uint64_t UNDEF = ~0ull;
constexpr int kLeafK = 4;
constexpr int kInternK = 16;
constexpr uint64_t kSnodSize = 8 + 2 * kLeafK * 40;
constexpr uint64_t kTreeSize = 24 + 2 * kInternK * 8 + (2 * kInternK + 1) * 8;
constexpr int kMaxSets = 2 * kInternK * 2 * kLeafK;   // one B-tree node of full nodes
constexpr uint64_t kFreeNull = 1;   // end of a local heap's free list

uint64_t pad8(uint64_t n) { return (n + 7) & ~7ull; }

struct Buf {
  std::vector<uint8_t> b;
  void u8(unsigned v) { b.push_back((uint8_t)v); }
  void u16(unsigned v) { u8(v & 0xff); u8((v >> 8) & 0xff); }
  void u32(uint32_t v) { for (int i = 0; i < 4; ++i) u8((v >> (8 * i)) & 0xff); }
  void u64(uint64_t v) { for (int i = 0; i < 8; ++i) u8((unsigned)((v >> (8 * i)) & 0xff)); }
  void raw(const void *p, size_t n) {
    const uint8_t *c = static_cast<const uint8_t *>(p);
    b.insert(b.end(), c, c + n);
  }
  void zeros(size_t n) { b.insert(b.end(), n, 0); }
  void align8() { zeros(pad8(b.size()) - b.size()); }
};

uint64_t get(const uint8_t *p, int n) {
  uint64_t v = 0;
  for (int i = n - 1; i >= 0; --i) v = (v << 8) | p[i];
  return v;
}

// ---- datatypes -----------------------------------------------------------
// members of the particle records (SavedParticle / SavedIon, particles.h:111-153,
// as registered with H5Tinsert at save.cpp:225-251)
struct Field {
  const char *name;
  uint32_t off;
  int kind;
};
const std::vector<Field> &fields_of(int kind) {
  static const std::vector<Field> part = {
      {"cell", 0, I32},  {"x", 8, F64},   {"y", 16, F64},      {"z", 24, F64}, {"px", 32, F64},
      {"py", 40, F64},   {"pz", 48, F64}, {"weight", 56, F64}, {"q2m", 64, F64}};
  static const std::vector<Field> ion = [] {
    std::vector<Field> v = part;
    v.push_back({"i_z", 72, I32});
    v.push_back({"sort", 76, I32});
    return v;
  }();
  return kind == SAVED_ION ? ion : part;
}

void enc_f64(Buf &o) {
  o.u8(0x11);                          // version 1, class 1 (floating point)
  o.u8(0x20); o.u8(63); o.u8(0);       // little-endian, implied mantissa MSB; sign bit 63
  o.u32(8);
  o.u16(0); o.u16(64);                 // bit offset, precision
  o.u8(52); o.u8(11); o.u8(0); o.u8(52);   // exponent location, size; mantissa location, size
  o.u32(1023);                         // exponent bias
}
void enc_int(Buf &o, uint32_t size) {
  o.u8(0x10);                          // version 1, class 0 (fixed point)
  o.u8(0x08); o.u8(0); o.u8(0);        // little-endian, signed
  o.u32(size);
  o.u16(0); o.u16((unsigned)(8 * size));
}
void enc_type(Buf &o, int kind) {
  if (kind == F64) return enc_f64(o);
  if (kind == I32) return enc_int(o, 4);
  if (kind == I64) return enc_int(o, 8);
  const std::vector<Field> &fs = fields_of(kind);
  o.u8(0x16);                          // version 1, class 6 (compound)
  o.u16((unsigned)fs.size()); o.u8(0);
  o.u32((uint32_t)kind_size(kind));
  for (const Field &f : fs) {
    const size_t n = strlen(f.name) + 1;
    o.raw(f.name, n);
    o.zeros(pad8(n) - n);
    o.u32(f.off);
    o.u8(0); o.zeros(3);               // scalar member
    o.u32(0); o.u32(0);                // permutation, reserved
    o.zeros(16);                       // dimension sizes
    if (f.kind == F64) enc_f64(o); else enc_int(o, 4);
  }
}

void message(Buf &o, unsigned type, const Buf &data, unsigned flags) {
  const uint64_t n = pad8(data.b.size());
  o.u16(type); o.u16((unsigned)n); o.u8(flags); o.zeros(3);
  o.raw(data.b.data(), data.b.size());
  o.zeros(n - data.b.size());
}

Buf object_header(const std::vector<std::pair<unsigned, Buf>> &msgs,
                  const std::vector<unsigned> &flags) {
  Buf body;
  for (size_t i = 0; i < msgs.size(); ++i) message(body, msgs[i].first, msgs[i].second, flags[i]);
  Buf o;
  o.u8(1); o.u8(0); o.u16((unsigned)msgs.size());
  o.u32(1);                            // reference count
  o.u32((uint32_t)body.b.size());
  o.zeros(4);                          // prefix aligned to 8
  o.raw(body.b.data(), body.b.size());
  return o;
}

Buf dataset_header(int kind, const std::vector<uint64_t> &dims, uint64_t addr, uint64_t bytes) {
  Buf space, type, fill, layout;
  space.u8(1); space.u8((unsigned)dims.size()); space.u8(0); space.u8(0); space.u32(0);
  for (uint64_t d : dims) space.u64(d);
  enc_type(type, kind);
  fill.u8(2); fill.u8(2); fill.u8(2); fill.u8(1); fill.u32(0);   // late alloc, default fill
  layout.u8(3); layout.u8(1); layout.u64(addr); layout.u64(bytes);
  return object_header({{0x0001, space}, {0x0003, type}, {0x0005, fill}, {0x0008, layout}},
                       {0, 1, 1, 0});
}

}  // namespace

size_t kind_size(int kind) {
  switch (kind) {
    case F64: return 8;
    case I32: return 4;
    case I64: return 8;
    case SAVED_PARTICLE: return 72;
    case SAVED_ION: return 80;
    default: return 0;
  }
}

uint64_t Dataset::count() const {
  uint64_t n = 1;
  for (uint64_t d : dims) n *= d;
  return n;
}

// ---- writer ----------------------------------------------------------------
Writer::~Writer() {
  if (f_) fclose(f_);
}

bool Writer::fail(const std::string &m) {
  err_ = m;
  if (f_) { fclose(f_); f_ = nullptr; }
  return false;
}

bool Writer::add(const std::string &name, int kind, const std::vector<uint64_t> &dims) {
  if (f_) return fail("add after open");
  if (name.empty() || name.find('/') != std::string::npos || name.find('\0') != std::string::npos)
    return fail("dataset name '" + name + "' must be non-empty without '/'");
  if (kind < 0 || kind >= NKIND) return fail("unknown element kind");
  if (dims.empty() || dims.size() > 3) return fail("rank must be 1..3");
  for (const Decl &d : sets_)
    if (d.name == name) return fail("duplicate dataset name '" + name + "'");
  if ((int)sets_.size() >= kMaxSets) return fail("too many datasets in one file");
  uint64_t n = kind_size(kind);
  for (uint64_t d : dims) n *= d;
  sets_.push_back({name, kind, dims, n, UNDEF, 0, 0});
  return true;
}

bool Writer::open(const std::string &path) {
  if (sets_.empty()) return fail("no datasets declared");
  std::sort(sets_.begin(), sets_.end(),
            [](const Decl &a, const Decl &b) { return strcmp(a.name.c_str(), b.name.c_str()) < 0; });
  // local heap: "" at 0, the names, one 16-byte free block at the end
  Buf heap;
  heap.zeros(8);
  for (Decl &d : sets_) {
    d.heap_off = heap.b.size();
    heap.raw(d.name.c_str(), d.name.size() + 1);
    heap.align8();
  }
  const uint64_t free_off = heap.b.size();
  heap.u64(kFreeNull);
  heap.u64(16);
  const int nsets = (int)sets_.size();
  const int nsnod = (nsets + 2 * kLeafK - 1) / (2 * kLeafK);

  const uint64_t a_root = 96, a_heap = a_root + 40, a_heapdata = a_heap + 32;
  const uint64_t a_tree = a_heapdata + heap.b.size();
  const uint64_t a_snod = a_tree + kTreeSize;
  uint64_t cur = a_snod + nsnod * kSnodSize;
  std::vector<uint64_t> oh_size(nsets);
  for (int i = 0; i < nsets; ++i) {
    sets_[i].oh_addr = cur;
    oh_size[i] = dataset_header(sets_[i].kind, sets_[i].dims, 0, 0).b.size();
    cur += oh_size[i];
  }
  for (Decl &d : sets_) {
    cur = pad8(cur);
    d.addr = d.bytes ? cur : UNDEF;
    cur += d.bytes;
  }
  const uint64_t eof = cur;

  Buf o;
  // superblock version 0
  const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  o.raw(sig, 8);
  o.u8(0); o.u8(0); o.u8(0); o.u8(0); o.u8(0); o.u8(8); o.u8(8); o.u8(0);
  o.u16(kLeafK); o.u16(kInternK);
  o.u32(0);                  // consistency flags
  o.u64(0);                  // base address
  o.u64(UNDEF);              // free-space info
  o.u64(eof);
  o.u64(UNDEF);              // driver information block
  o.u64(0); o.u64(a_root); o.u32(1); o.u32(0); o.u64(a_tree); o.u64(a_heap);   // root entry
  // root group: one symbol table message
  {
    Buf stab;
    stab.u64(a_tree); stab.u64(a_heap);
    Buf h = object_header({{0x0011, stab}}, {0});
    o.raw(h.b.data(), h.b.size());
  }
  // local heap
  o.raw("HEAP", 4); o.u8(0); o.zeros(3);
  o.u64(heap.b.size()); o.u64(free_off); o.u64(a_heapdata);
  o.raw(heap.b.data(), heap.b.size());
  // B-tree node: key0 = "" then, per child, the last name of that child
  o.raw("TREE", 4); o.u8(0); o.u8(0); o.u16((unsigned)nsnod);
  o.u64(UNDEF); o.u64(UNDEF);
  o.u64(0);
  for (int s = 0; s < nsnod; ++s) {
    const int last = std::min(nsets, (s + 1) * 2 * kLeafK) - 1;
    o.u64(a_snod + s * kSnodSize);
    o.u64(sets_[last].heap_off);
  }
  if (o.b.size() > a_snod) return fail("internal layout error (B-tree)");
  o.zeros(a_snod - o.b.size());
  // symbol-table nodes
  for (int s = 0; s < nsnod; ++s) {
    const int i0 = s * 2 * kLeafK, i1 = std::min(nsets, i0 + 2 * kLeafK);
    const size_t start = o.b.size();
    o.raw("SNOD", 4); o.u8(1); o.u8(0); o.u16((unsigned)(i1 - i0));
    for (int i = i0; i < i1; ++i) {
      o.u64(sets_[i].heap_off); o.u64(sets_[i].oh_addr); o.u32(0); o.u32(0); o.zeros(16);
    }
    o.zeros(kSnodSize - (o.b.size() - start));
  }
  for (int i = 0; i < nsets; ++i) {
    Buf h = dataset_header(sets_[i].kind, sets_[i].dims, sets_[i].addr, sets_[i].bytes);
    if (o.b.size() != sets_[i].oh_addr || h.b.size() != oh_size[i])
      return fail("internal layout error (object header)");
    o.raw(h.b.data(), h.b.size());
  }
  f_ = fopen(path.c_str(), "wb");
  if (!f_) return fail("cannot create " + path + ": " + strerror(errno));
  if (fwrite(o.b.data(), 1, o.b.size(), f_) != o.b.size()) return fail("write failed: " + path);
  // extend to the end-of-file address; the data regions are filled by put()
  if (eof > o.b.size()) {
    if (fseeko(f_, (off_t)(eof - 1), SEEK_SET) != 0 || fputc(0, f_) == EOF)
      return fail("cannot extend " + path);
  }
  return true;
}

bool Writer::put(int i, uint64_t at, const void *src, uint64_t bytes) {
  if (!f_) return fail("put before open");
  if (i < 0 || i >= (int)sets_.size()) return fail("dataset index out of range");
  const Decl &d = sets_[i];
  if (at > d.bytes || bytes > d.bytes - at) return fail("write past the end of '" + d.name + "'");
  if (!bytes) return true;
  if (fseeko(f_, (off_t)(d.addr + at), SEEK_SET) != 0 ||
      fwrite(src, 1, bytes, f_) != bytes)
    return fail("write failed for '" + d.name + "'");
  return true;
}

bool Writer::close() {
  if (!f_) return err_.empty() ? fail("close before open") : false;
  const bool ok = fflush(f_) == 0;
  fclose(f_);
  f_ = nullptr;
  return ok ? true : fail("flush failed");
}

// ---- reader ----------------------------------------------------------------
Reader::~Reader() {
  if (f_) fclose(f_);
}

bool Reader::fail(const std::string &m) {
  err_ = m;
  return false;
}

bool Reader::at(uint64_t addr, void *dst, uint64_t n) {
  if (addr == UNDEF || addr > eof_ || n > eof_ - addr)
    return fail("address out of the file (truncated or damaged file)");
  if (fseeko(f_, (off_t)addr, SEEK_SET) != 0 || fread(dst, 1, n, f_) != n)
    return fail("read failed");
  return true;
}

namespace {

// decodes a datatype message at p (n bytes); returns bytes used, 0 on error
size_t dec_type(const uint8_t *p, size_t n, Type &t, std::string &err) {
  if (n < 8) { err = "truncated datatype"; return 0; }
  const int version = p[0] >> 4, cls = p[0] & 0x0f;
  const uint32_t bits = (uint32_t)get(p + 1, 3);
  t = Type();
  t.cls = cls;
  t.size = (uint32_t)get(p + 4, 4);
  size_t used = 8;
  switch (cls) {
    case 0:   // fixed point
      t.big_endian = bits & 1;
      t.is_signed = (bits >> 3) & 1;
      used += 4;
      break;
    case 1:   // floating point
      if (bits & 0x40) { err = "VAX floating point is not supported"; return 0; }
      t.big_endian = bits & 1;
      used += 12;
      break;
    case 3:   // string: no properties
      break;
    case 4:   // bitfield
      t.big_endian = bits & 1;
      used += 4;
      break;
    case 6: {  // compound
      const int nm = (int)(bits & 0xffff);
      for (int m = 0; m < nm; ++m) {
        Member mb;
        const uint8_t *q = p + used;
        size_t len = 0;
        while (used + len < n && q[len]) ++len;
        if (used + len >= n) { err = "truncated compound member name"; return 0; }
        mb.name.assign(reinterpret_cast<const char *>(q), len);
        used += version >= 3 ? len + 1 : pad8(len + 1);
        if (version >= 3) {
          int ob = 1;
          while (ob < 4 && (t.size >> (8 * ob))) ++ob;
          if (used + ob > n) { err = "truncated compound member"; return 0; }
          mb.offset = (uint32_t)get(p + used, ob);
          used += ob;
        } else {
          if (used + 4 > n) { err = "truncated compound member"; return 0; }
          mb.offset = (uint32_t)get(p + used, 4);
          used += 4;
          if (version == 1) {
            if (used + 28 > n) { err = "truncated compound member"; return 0; }
            if (p[used] != 0) { err = "array members of compounds are not supported"; return 0; }
            used += 28;
          }
        }
        const size_t u = dec_type(p + used, n - used, mb.type, err);
        if (!u) return 0;
        used += u;
        t.members.push_back(mb);
      }
      break;
    }
    default:
      err = "datatype class " + std::to_string(cls) + " is not supported";
      return 0;
  }
  if (used > n) { err = "truncated datatype"; return 0; }
  return used;
}

}  // namespace

bool Reader::open(const std::string &path) {
  f_ = fopen(path.c_str(), "rb");
  if (!f_) return fail("cannot open " + path + ": " + strerror(errno));
  struct stat st;
  if (fstat(fileno(f_), &st) != 0) return fail("cannot stat " + path);
  eof_ = (uint64_t)st.st_size;
  uint8_t sb[100];
  if (!at(0, sb, 96)) return fail(path + ": not an HDF5 file (too short)");
  const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  if (memcmp(sb, sig, 8) != 0) return fail(path + ": not an HDF5 file (no signature at 0)");
  const int version = sb[8];
  if (version > 1)
    return fail(path + ": superblock version " + std::to_string(version) +
                " (newer-format files) is not supported");
  if (sb[13] != 8 || sb[14] != 8) return fail(path + ": only 8-byte offsets and lengths are supported");
  const int off = version == 0 ? 24 : 28;
  if (version == 1 && !at(0, sb, 100)) return false;
  const uint64_t base = get(sb + off, 8);
  if (base != 0) return fail(path + ": a non-zero base address is not supported");
  const uint8_t *root = sb + off + 32;
  const uint64_t root_oh = get(root + 8, 8);
  // root object header: the symbol table message
  uint8_t pre[16];
  if (!at(root_oh, pre, 16)) return false;
  if (pre[0] != 1) return fail(path + ": root object header version is not 1");
  const uint64_t hsize = get(pre + 8, 4);
  std::vector<uint8_t> hb(hsize);
  if (!at(root_oh + 16, hb.data(), hsize)) return false;
  uint64_t tree = UNDEF, heap = UNDEF;
  for (size_t q = 0; q + 8 <= hsize;) {
    const unsigned type = (unsigned)get(&hb[q], 2), sz = (unsigned)get(&hb[q + 2], 2);
    if (q + 8 + sz > hsize) break;
    if (type == 0x11 && sz >= 16) {
      tree = get(&hb[q + 8], 8);
      heap = get(&hb[q + 16], 8);
    }
    q += 8 + sz;
  }
  if (tree == UNDEF && get(root + 16, 4) == 1) {   // cached in the root entry
    tree = get(root + 24, 8);
    heap = get(root + 32, 8);
  }
  if (tree == UNDEF) return fail(path + ": the root group is not a symbol-table group");
  uint8_t hh[32];
  if (!at(heap, hh, 32)) return false;
  if (memcmp(hh, "HEAP", 4) != 0) return fail(path + ": bad local heap");
  const uint64_t heap_size = get(hh + 8, 8), heap_data = get(hh + 24, 8);
  if (!walk_btree(tree, heap_data, heap_size, 0)) return fail(path + ": " + err_);
  return true;
}

bool Reader::walk_btree(uint64_t addr, uint64_t heap_data, uint64_t heap_size, int depth) {
  if (depth > 32) return fail("group B-tree too deep (cycle?)");
  uint8_t h[24];
  if (!at(addr, h, 24)) return false;
  if (memcmp(h, "TREE", 4) != 0 || h[4] != 0) return fail("bad group B-tree node");
  const int level = h[5], used = (int)get(h + 6, 2);
  std::vector<uint8_t> kc(16 * (size_t)used + 8);
  if (!at(addr + 24, kc.data(), kc.size())) return false;
  for (int c = 0; c < used; ++c) {
    const uint64_t child = get(&kc[8 + 16 * (size_t)c], 8);
    if (level > 0) {
      if (!walk_btree(child, heap_data, heap_size, depth + 1)) return false;
      continue;
    }
    uint8_t sh[8];
    if (!at(child, sh, 8)) return false;
    if (memcmp(sh, "SNOD", 4) != 0) return fail("bad symbol-table node");
    const int nsym = (int)get(sh + 6, 2);
    std::vector<uint8_t> ent(40 * (size_t)nsym);
    if (!at(child + 8, ent.data(), ent.size())) return false;
    for (int e = 0; e < nsym; ++e) {
      const uint8_t *en = &ent[40 * (size_t)e];
      const uint64_t name_off = get(en, 8), oh = get(en + 8, 8);
      if (name_off >= heap_size) return fail("symbol name outside the local heap");
      std::string name;
      const uint64_t room = std::min<uint64_t>(heap_size - name_off, 4096);
      std::vector<char> nb(room);
      if (!at(heap_data + name_off, nb.data(), room)) return false;
      const void *z = memchr(nb.data(), 0, room);
      if (!z) return fail("unterminated symbol name");
      name.assign(nb.data(), static_cast<const char *>(z) - nb.data());
      if (get(en + 16, 4) == 1) continue;   // a sub-group
      if (!read_object(name, oh)) return false;
    }
  }
  return true;
}

bool Reader::read_object(const std::string &name, uint64_t oh) {
  uint8_t pre[16];
  if (!at(oh, pre, 16)) return false;
  if (memcmp(pre, "OHDR", 4) == 0)
    return fail("'" + name + "': version-2 object headers are not supported");
  if (pre[0] != 1) return fail("'" + name + "': unknown object header version");
  const unsigned nmsg = (unsigned)get(pre + 2, 2);
  std::vector<std::pair<uint64_t, uint64_t>> chunks = {{oh + 16, get(pre + 8, 4)}};
  Dataset d;
  d.name = name;
  bool have_space = false, have_type = false, have_layout = false;
  unsigned seen = 0;
  for (size_t ci = 0; ci < chunks.size() && seen < nmsg; ++ci) {
    if (ci > 64) return fail("'" + name + "': too many header continuation blocks");
    std::vector<uint8_t> b(chunks[ci].second);
    if (!at(chunks[ci].first, b.data(), b.size())) return false;
    for (size_t q = 0; q + 8 <= b.size() && seen < nmsg;) {
      const unsigned type = (unsigned)get(&b[q], 2), sz = (unsigned)get(&b[q + 2], 2);
      const unsigned flags = b[q + 4];
      if (q + 8 + sz > b.size()) return fail("'" + name + "': truncated header message");
      const uint8_t *m = &b[q + 8];
      ++seen;
      if ((flags & 2) && (type == 0x0003 || type == 0x0001))
        return fail("'" + name + "': shared (committed) datatypes are not supported");
      if (type == 0x0001) {   // dataspace
        const int ver = m[0], rank = m[1];
        const size_t d0 = ver == 1 ? 8 : 4;
        if (d0 + 8 * (size_t)rank > sz) return fail("'" + name + "': bad dataspace");
        d.dims.clear();
        for (int r = 0; r < rank; ++r) d.dims.push_back(get(m + d0 + 8 * r, 8));
        if (ver == 2 && m[3] == 2) d.dims.assign(1, 0);   // null dataspace
        have_space = true;
      } else if (type == 0x0003) {
        std::string e;
        if (!dec_type(m, sz, d.type, e)) return fail("'" + name + "': " + e);
        have_type = true;
      } else if (type == 0x0008) {   // layout
        const int ver = m[0];
        if (ver == 3 || ver == 4) {
          const int cls = m[1];
          if (cls == 0) {
            const uint64_t n = get(m + 2, 2);
            if (4 + n > sz) return fail("'" + name + "': bad compact layout");
            d.compact = true;
            d.inline_data.assign(m + 4, m + 4 + n);
            d.bytes = n;
          } else if (cls == 1) {
            d.addr = get(m + 2, 8);
            d.bytes = get(m + 10, 8);
          } else {
            return fail("'" + name + "': chunked or virtual storage is not supported");
          }
        } else if (ver == 1 || ver == 2) {
          const int nd = m[1], cls = m[2];
          size_t q2 = 8;
          if (cls == 2) return fail("'" + name + "': chunked storage is not supported");
          if (cls == 1) { d.addr = get(m + q2, 8); q2 += 8; }
          uint64_t n = 1;
          for (int r = 0; r < nd; ++r) n *= get(m + q2 + 4 * r, 4);
          q2 += 4 * (size_t)nd;
          if (cls == 0) {
            const uint64_t cn = get(m + q2, 4);
            if (q2 + 4 + cn > sz) return fail("'" + name + "': bad compact layout");
            d.compact = true;
            d.inline_data.assign(m + q2 + 4, m + q2 + 4 + cn);
            n = cn;
          }
          d.bytes = n;
        } else {
          return fail("'" + name + "': layout message version " + std::to_string(ver));
        }
        have_layout = true;
      } else if (type == 0x0010) {   // continuation
        chunks.push_back({get(m, 8), get(m + 8, 8)});
      }
      q += 8 + sz;
    }
  }
  if (!(have_space && have_type && have_layout)) return true;   // not a dataset
  if (!d.compact && d.addr != UNDEF && (d.addr > eof_ || d.bytes > eof_ - d.addr))
    return fail("'" + name + "': data outside the file (truncated file)");
  const uint64_t need = d.count() * d.type.size;
  if (d.bytes < need && (d.compact || d.addr != UNDEF))
    return fail("'" + name + "': storage smaller than its dataspace");
  sets_.push_back(std::move(d));
  return true;
}

const Dataset *Reader::find(const std::string &name) const {
  for (const Dataset &d : sets_)
    if (d.name == name) return &d;
  return nullptr;
}

bool Reader::raw(const Dataset &d, uint64_t at0, void *dst, uint64_t bytes) {
  const uint64_t total = d.count() * d.type.size;
  if (at0 > total || bytes > total - at0) return fail("read past the end of '" + d.name + "'");
  if (d.compact) {
    memcpy(dst, d.inline_data.data() + at0, bytes);
    return true;
  }
  if (d.addr == UNDEF) {
    memset(dst, 0, bytes);
    return true;
  }
  return at(d.addr + at0, dst, bytes);
}

namespace {

bool scalar(const uint8_t *src, const Type &t, int want, double &dv, int64_t &iv) {
  uint8_t b[8];
  if (t.size > 8 || t.size == 0) return false;
  for (uint32_t i = 0; i < t.size; ++i) b[i] = t.big_endian ? src[t.size - 1 - i] : src[i];
  if (t.cls == 1) {
    if (want != F64) return false;
    if (t.size == 8) { memcpy(&dv, b, 8); return true; }
    if (t.size == 4) { float f; memcpy(&f, b, 4); dv = f; return true; }
    return false;
  }
  if (t.cls != 0) return false;
  uint64_t u = get(b, (int)t.size);
  if (t.is_signed && t.size < 8 && (u >> (8 * t.size - 1)) & 1) u |= ~0ull << (8 * t.size);
  iv = (int64_t)u;
  dv = t.is_signed ? (double)iv : (double)u;
  return true;
}

}  // namespace

bool Reader::read(const Dataset &d, int kind, uint64_t first, uint64_t count, void *dst) {
  if (kind < 0 || kind >= NKIND) return fail("unknown element kind");
  const uint64_t n = d.count();
  if (first > n || count > n - first) return fail("element range outside '" + d.name + "'");
  const uint32_t es = d.type.size;
  const size_t ks = kind_size(kind);
  // compound target: source member of every target field
  std::vector<const Member *> src_m;
  if (kind == SAVED_PARTICLE || kind == SAVED_ION) {
    if (d.type.cls != 6) return fail("'" + d.name + "' is not a compound dataset");
    for (const Field &f : fields_of(kind)) {
      const Member *m = nullptr;
      for (const Member &mm : d.type.members)
        if (mm.name == f.name) m = &mm;
      if (!m) return fail("'" + d.name + "' has no member '" + f.name + "'");
      if (m->offset + m->type.size > es) return fail("'" + d.name + "': bad member offset");
      src_m.push_back(m);
    }
  } else if (d.type.cls != 0 && d.type.cls != 1) {
    return fail("'" + d.name + "' is not a numeric dataset");
  } else if (d.type.cls == 1 && kind != F64) {
    return fail("'" + d.name + "' holds floating-point values, not integers");
  }
  const uint64_t block = std::max<uint64_t>(1, (8u << 20) / std::max<uint32_t>(es, 1));
  std::vector<uint8_t> buf;
  uint8_t *out = static_cast<uint8_t *>(dst);
  for (uint64_t e0 = 0; e0 < count; e0 += block) {
    const uint64_t ne = std::min(block, count - e0);
    buf.resize(ne * es);
    if (!raw(d, (first + e0) * es, buf.data(), ne * es)) return false;
    for (uint64_t e = 0; e < ne; ++e) {
      const uint8_t *s = &buf[e * es];
      uint8_t *o = out + (e0 + e) * ks;
      double dv = 0;
      int64_t iv = 0;
      if (src_m.empty()) {
        if (!scalar(s, d.type, kind == F64 ? F64 : I64, dv, iv))
          return fail("'" + d.name + "': element type cannot be converted");
        if (kind == F64) memcpy(o, &dv, 8);
        else if (kind == I64) memcpy(o, &iv, 8);
        else { const int32_t v = (int32_t)iv; memcpy(o, &v, 4); }
        continue;
      }
      memset(o, 0, ks);
      const std::vector<Field> &fs = fields_of(kind);
      for (size_t k = 0; k < fs.size(); ++k) {
        if (!scalar(s + src_m[k]->offset, src_m[k]->type, fs[k].kind, dv, iv))
          return fail("'" + d.name + "': member '" + fs[k].name + "' cannot be converted");
        if (fs[k].kind == F64) memcpy(o + fs[k].off, &dv, 8);
        else { const int32_t v = (int32_t)iv; memcpy(o + fs[k].off, &v, 4); }
      }
    }
  }
  return true;
}

}  // namespace h5

// ---- C-ABI ---------------------------------------------------------------------
namespace {
int h5_fail(const std::string &m) {
  qss_set_error(m);
  return QSS_E_ARG;
}
int kind_of(const h5::Dataset &d) {
  if (d.type.cls == 1) return h5::F64;
  if (d.type.cls == 0) return d.type.size <= 4 ? h5::I32 : h5::I64;
  if (d.type.cls == 6) {
    bool ion = false, sort = false;
    for (const h5::Member &m : d.type.members) {
      ion |= m.name == "i_z";
      sort |= m.name == "sort";
    }
    return ion && sort ? h5::SAVED_ION : h5::SAVED_PARTICLE;
  }
  return -1;
}
}  // namespace

extern "C" {

int qss_h5_write(const char *path, int n, const qss_h5_set *sets) {
  if (!path || n <= 0 || !sets) return h5_fail("h5_write: path and at least one set required");
  h5::Writer w;
  for (int i = 0; i < n; ++i) {
    const qss_h5_set &s = sets[i];
    if (!s.name || s.rank < 1 || s.rank > 3) return h5_fail("h5_write: set needs a name and rank 1..3");
    std::vector<uint64_t> dims;
    for (int r = 0; r < s.rank; ++r) {
      if (s.dims[r] < 0) return h5_fail("h5_write: negative dimension");
      dims.push_back((uint64_t)s.dims[r]);
    }
    if (!w.add(s.name, s.kind, dims)) return h5_fail("h5_write: " + w.error());
  }
  if (!w.open(path)) return h5_fail("h5_write: " + w.error());
  // the writer sorted its sets by name: match them back to the caller's
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return strcmp(sets[a].name, sets[b].name) < 0; });
  for (int i = 0; i < n; ++i) {
    const qss_h5_set &s = sets[order[i]];
    uint64_t bytes = h5::kind_size(s.kind);
    for (int r = 0; r < s.rank; ++r) bytes *= (uint64_t)s.dims[r];
    if (bytes && !s.data) return h5_fail(std::string("h5_write: no data for ") + s.name);
    if (!w.put(i, 0, s.data, bytes)) return h5_fail("h5_write: " + w.error());
  }
  if (!w.close()) return h5_fail("h5_write: " + w.error());
  return QSS_OK;
}

int qss_h5_info(const char *path, const char *name, int32_t *kind, int32_t *rank, int64_t *dims) {
  if (!path || !name) return h5_fail("h5_info: path and name required");
  h5::Reader r;
  if (!r.open(path)) return h5_fail("h5_info: " + r.error());
  const h5::Dataset *d = r.find(name);
  if (!d) return h5_fail(std::string("h5_info: no dataset '") + name + "' in " + path);
  if (kind) *kind = kind_of(*d);
  if (rank) *rank = (int32_t)d->dims.size();
  if (dims)
    for (size_t i = 0; i < 3; ++i) dims[i] = i < d->dims.size() ? (int64_t)d->dims[i] : 0;
  return QSS_OK;
}

int qss_h5_read(const char *path, const char *name, int32_t kind, int64_t first, int64_t count,
                void *dst) {
  if (!path || !name || first < 0 || count < 0 || (count > 0 && !dst))
    return h5_fail("h5_read: bad arguments");
  h5::Reader r;
  if (!r.open(path)) return h5_fail("h5_read: " + r.error());
  const h5::Dataset *d = r.find(name);
  if (!d) return h5_fail(std::string("h5_read: no dataset '") + name + "' in " + path);
  if (!r.read(*d, kind, (uint64_t)first, (uint64_t)count, dst))
    return h5_fail("h5_read: " + r.error());
  return QSS_OK;
}

int qss_h5_list(const char *path, char *names, size_t cap) {
  if (!path || (cap && !names)) return h5_fail("h5_list: bad arguments");
  h5::Reader r;
  if (!r.open(path)) return h5_fail("h5_list: " + r.error());
  std::string all;
  for (const h5::Dataset &d : r.datasets()) all += d.name + "\n";
  if (all.size() + 1 > cap) return h5_fail("h5_list: buffer of " + std::to_string(cap) +
                                           " bytes too small, need " + std::to_string(all.size() + 1));
  memcpy(names, all.c_str(), all.size() + 1);
  return QSS_OK;
}

}  // extern "C"
