// h5file.h — the subset of the HDF5 file format that the reference's output
// files use, written and read without libhdf5.
//
// Mesh::Save (save.cpp:178-835) and its record helpers (myhdfshell.cpp:
// WriteHDF5Recorddouble / RecordInt / RecordLong) create files with the
// library defaults: superblock version 0, one root group kept as a symbol
// table (version-1 B-tree of symbol-table nodes, names in a local heap),
// version-1 object headers, and contiguous datasets of IEEE F64LE, STD I32LE,
// STD I64LE or a compound of those (the SavedParticle / SavedIon records,
// save.cpp:225-251).  Writer produces exactly that structure; Reader parses it
// (also several B-tree levels, header continuation blocks, compact layout and
// older layout-message versions, which files of other writers of this format
// version contain) and refuses anything else with a message.
//
// Writer::open lays out every piece up front: metadata first, then the raw
// data of each dataset at a fixed address, so that large datasets are
// streamed (put) in any number of pieces without being held in memory.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace h5 {

enum Kind { F64 = 0, I32 = 1, I64 = 2, SAVED_PARTICLE = 3, SAVED_ION = 4, NKIND = 5 };

// size in bytes of one element of a kind (the C records of qss.h)
size_t kind_size(int kind);

// a scalar or compound datatype as stored in the file
struct Member;
struct Type {
  int cls = -1;            // 0 fixed-point, 1 floating-point, 6 compound
  uint32_t size = 0;
  bool is_signed = false;  // fixed-point
  bool big_endian = false;
  std::vector<Member> members;   // compound
};
struct Member {
  std::string name;
  uint32_t offset = 0;
  Type type;
};

struct Dataset {
  std::string name;
  std::vector<uint64_t> dims;
  Type type;
  bool compact = false;          // raw data kept in the header
  std::vector<uint8_t> inline_data;
  uint64_t addr = ~0ull;         // contiguous data; ~0 = never written
  uint64_t bytes = 0;
  uint64_t count() const;
};

class Writer {
 public:
  ~Writer();
  // declare every dataset (unique names, 1 <= rank <= 3) before open()
  bool add(const std::string &name, int kind, const std::vector<uint64_t> &dims);
  bool open(const std::string &path);
  // writes `bytes` of dataset i's raw data starting at byte `at` of it
  bool put(int i, uint64_t at, const void *src, uint64_t bytes);
  bool close();
  const std::string &error() const { return err_; }

 private:
  struct Decl {
    std::string name;
    int kind;
    std::vector<uint64_t> dims;
    uint64_t bytes, addr, oh_addr, heap_off;
  };
  std::vector<Decl> sets_;
  FILE *f_ = nullptr;
  std::string err_;
  bool fail(const std::string &m);
};

class Reader {
 public:
  ~Reader();
  bool open(const std::string &path);
  const Dataset *find(const std::string &name) const;
  const std::vector<Dataset> &datasets() const { return sets_; }
  // raw bytes [at, at+bytes) of a dataset; unallocated data reads as zeros
  bool raw(const Dataset &d, uint64_t at, void *dst, uint64_t bytes);
  // reads `count` elements starting at element `first`, converted to `kind`
  // (numbers between the integer and float sizes; compounds member by member,
  // matched by name)
  bool read(const Dataset &d, int kind, uint64_t first, uint64_t count, void *dst);
  const std::string &error() const { return err_; }

 private:
  FILE *f_ = nullptr;
  uint64_t eof_ = 0;
  std::vector<Dataset> sets_;
  std::string err_;
  bool fail(const std::string &m);
  bool at(uint64_t addr, void *dst, uint64_t n);
  bool walk_btree(uint64_t addr, uint64_t heap_data, uint64_t heap_size, int depth);
  bool read_object(const std::string &name, uint64_t oh_addr);
};

}  // namespace h5
