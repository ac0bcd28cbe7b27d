// Sparse-grid substrate for the multi-evaluation path: grid points, hash
// storage, regular generation, surplus refinement and hierarchisation for the
// Linear and ModLinear (no-boundary) grid types.
//
// Sequence numbers must match SG++ exactly, because the operator's inputs
// (alpha, level/index arrays) are indexed by them. The pieces that fix them:
//   - the point hash (base/src/sgpp/base/grid/storage/hashmap/HashGridPoint.cpp:151-160)
//     and content equality (:164-178);
//   - a std::unordered_map keyed by point pointers with that hash, driven by
//     the same insert / erase sequence (HashGridStorage.cpp:288-292, 325-336),
//     so that its iteration order, which refinement depends on, is the same;
//   - regular_iter's insertion order (HashGenerator.hpp:282-342);
//   - refinement's top-k heap and ancestor creation
//     (HashRefinement.cpp:19-141, AbstractRefinement.cpp:36-52,
//      AbstractRefinement.hpp:229-244).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "sgpp/base/datatypes/DataVector.hpp"
#include "sgpp/base/exception/exceptions.hpp"

namespace sgpp {
namespace base {

typedef uint32_t level_t;
typedef uint32_t index_t;

class HashGridPoint {
 public:
  typedef uint32_t level_type;
  typedef uint32_t index_type;

  HashGridPoint() = default;
  explicit HashGridPoint(size_t dim) : level_(dim, 0), index_(dim, 0) {}
  HashGridPoint(const HashGridPoint& o) : level_(o.level_), index_(o.index_), leaf_(o.leaf_) {
    copy_hash(o);
  }
  HashGridPoint& operator=(const HashGridPoint& o) {
    if (this != &o) {
      level_ = o.level_;
      index_ = o.index_;
      leaf_ = o.leaf_;
      copy_hash(o);
    }
    return *this;
  }

  size_t getDimension() const { return level_.size(); }
  void set(size_t d, level_type l, index_type i) {
    const size_t before = term(d);
    level_[d] = l;
    index_[d] = i;
    update_hash(d, before);
  }
  void set(size_t d, level_type l, index_type i, bool isLeaf) {
    leaf_ = isLeaf;
    set(d, l, i);
  }
  // no rehash (HashGridPoint.hpp:125-143); callers copy the point afterwards
  void push(size_t d, level_type l, index_type i) {
    level_[d] = l;
    index_[d] = i;
    hash_valid_ = false;
  }
  void push(size_t d, level_type l, index_type i, bool isLeaf) {
    push(d, l, i);
    leaf_ = isLeaf;
  }
  void get(size_t d, level_type& l, index_type& i) const {
    l = level_[d];
    i = index_[d];
  }
  level_type getLevel(size_t d) const { return level_[d]; }
  index_type getIndex(size_t d) const { return index_[d]; }
  void setLeaf(bool b) { leaf_ = b; }
  bool isLeaf() const { return leaf_; }
  // Set by refinement on a stored point once all 2d children are in its
  // storage; points are never removed one by one, so it stays true. Not
  // copied: a copy may live in a storage without those children.
  bool childrenComplete() const { return children_complete_; }
  void setChildrenComplete() { children_complete_ = true; }
  size_t getHash() const { return hash_; }
  bool equals(const HashGridPoint& o) const {
    return level_ == o.level_ && index_ == o.index_;
  }
  level_type getLevelSum() const {
    level_type s = 0;
    for (level_type l : level_) s += l;
    return s;
  }
  level_type getLevelMax() const {
    level_type m = level_[0];
    for (level_type l : level_) m = l > m ? l : m;
    return m;
  }
  double getStandardCoordinate(size_t d) const {
    return static_cast<double>(index_[d]) / static_cast<double>(uint64_t(1) << level_[d]);
  }
  // HashGridPoint.cpp:151-178: h = 0xdeadbeef; h = (2^l_d + i_d) + h * 65599
  // over d = 0..dim-1 (the level-index sum in 32 bits, h in 64 bits)
  void rehash() {
    size_t h = 0xdeadbeef;
    for (size_t d = 0; d < level_.size(); ++d) h = term(d) + h * kHashMul;
    hash_ = h;
    hash_valid_ = true;
  }

 private:
  static constexpr size_t kHashMul = 65599;
  static constexpr size_t kMaxIncrementalDim = 64;
  // the hash is a polynomial in kHashMul, so changing the term of dimension
  // d changes it by (new - old) * kHashMul^(dim - 1 - d) modulo 2^64: set()
  // updates it in O(1) and yields the value rehash() would compute
  static size_t hash_power(size_t k) {
    static const std::vector<size_t> table = [] {
      std::vector<size_t> t(kMaxIncrementalDim, 1);
      for (size_t j = 1; j < t.size(); ++j) t[j] = t[j - 1] * kHashMul;
      return t;
    }();
    return table[k];
  }
  size_t term(size_t d) const {
    return static_cast<index_type>((static_cast<index_type>(1) << level_[d]) + index_[d]);
  }
  void update_hash(size_t d, size_t before) {
    if (hash_valid_ && level_.size() <= kMaxIncrementalDim)
      hash_ += (term(d) - before) * hash_power(level_.size() - 1 - d);
    else
      rehash();
  }
  void copy_hash(const HashGridPoint& o) {
    if (o.hash_valid_) {
      hash_ = o.hash_;
      hash_valid_ = true;
    } else {
      rehash();
    }
  }

  std::vector<level_type> level_;
  std::vector<index_type> index_;
  size_t hash_ = 0;
  bool hash_valid_ = false;
  bool leaf_ = false;
  bool children_complete_ = false;
};

// Not noexcept on purpose: matches the reference functors, which makes
// libstdc++ cache hash codes in the nodes exactly as it does there.
struct HashGridPointPointerHashFunctor {
  size_t operator()(const HashGridPoint* p) const { return p->getHash(); }
};
struct HashGridPointPointerEqualityFunctor {
  bool operator()(const HashGridPoint* a, const HashGridPoint* b) const { return a->equals(*b); }
};

class HashGridStorage {
 public:
  typedef HashGridPoint point_type;
  typedef std::unordered_map<HashGridPoint*, size_t, HashGridPointPointerHashFunctor,
                             HashGridPointPointerEqualityFunctor>
      grid_map;
  typedef grid_map::iterator grid_map_iterator;

  explicit HashGridStorage(size_t dim) : dim_(dim) {}
  ~HashGridStorage() { clear(); }
  HashGridStorage(const HashGridStorage&) = delete;
  HashGridStorage& operator=(const HashGridStorage&) = delete;

  size_t getSize() const { return map_.size(); }
  size_t getDimension() const { return dim_; }
  HashGridPoint& getPoint(size_t seq) { return *list_[seq]; }
  const HashGridPoint& getPoint(size_t seq) const { return *list_[seq]; }
  HashGridPoint& operator[](size_t seq) { return *list_[seq]; }

  size_t insert(const HashGridPoint& p);            // HashGridStorage.cpp:288-292
  void update(const HashGridPoint& p, size_t pos);  // HashGridStorage.cpp:325-336
  bool isContaining(const HashGridPoint& p) const {
    return map_.find(const_cast<HashGridPoint*>(&p)) != map_.end();
  }
  grid_map_iterator find(HashGridPoint* p) { return map_.find(p); }
  grid_map_iterator begin() { return map_.begin(); }
  grid_map_iterator end() { return map_.end(); }
  size_t getSequenceNumber(const HashGridPoint& p) const {
    auto it = map_.find(const_cast<HashGridPoint*>(&p));
    return it == map_.end() ? static_cast<size_t>(-1) : it->second;
  }
  bool isInvalidSequenceNumber(size_t s) const { return s >= list_.size(); }
  void clear();

  // N x d row-major level and index arrays in sequence order
  void getLevelIndexArrays(std::vector<uint32_t>& level, std::vector<uint32_t>& index) const;
  // 2^l and i as FP64 (HashGridStorage.cpp:433-450)
  void getLevelIndexArraysForEval(std::vector<double>& level, std::vector<double>& index) const;
  // (level, index, mask, offset) encoding of the ModLinear streaming kernel
  // (HashGridStorage.cpp:519-560)
  void getLevelIndexMaskArraysForModEval(std::vector<double>& level, std::vector<double>& index,
                                         std::vector<uint64_t>& mask,
                                         std::vector<double>& offset) const;
  // point coordinates, N x d row-major
  void getCoordinateArrays(std::vector<double>& coords) const;
  size_t getMaxLevel() const;
  // text form of HashGridStorage::serialize (HashGridStorage.cpp:217-243),
  // trivial bounding box, format version 9; unserialize reads versions 1-9
  std::string serialize() const;
  void unserialize(const std::string& text);
  void recalcLeafProperty();  // HashGridStorage.cpp:358-397

 private:
  size_t dim_;
  std::vector<HashGridPoint*> list_;
  grid_map map_;
};

enum class GridType { Linear = 0, ModLinear = 1 };

class RefinementFunctor {
 public:
  virtual ~RefinementFunctor() {}
  virtual double operator()(HashGridStorage& storage, size_t seq) const = 0;
  virtual size_t getRefinementsNum() const = 0;
  virtual double getRefinementThreshold() const = 0;
};

// |alpha_seq| (base/src/sgpp/base/grid/generation/functors/SurplusRefinementFunctor.cpp:23-27)
class SurplusRefinementFunctor : public RefinementFunctor {
 public:
  SurplusRefinementFunctor(const DataVector& alpha, size_t refinements_num = 1,
                           double threshold = 0.0)
      : alpha_(alpha), num_(refinements_num), thr_(threshold) {}
  double operator()(HashGridStorage&, size_t seq) const override;
  size_t getRefinementsNum() const override { return num_; }
  double getRefinementThreshold() const override { return thr_; }

 private:
  const DataVector& alpha_;
  size_t num_;
  double thr_;
};

// Highest level refinement may create: refine() throws generation_exception
// instead of refining a point that already has this level in some dimension.
constexpr level_t kMaxRefineLevel = 31;

class GridGenerator {
 public:
  explicit GridGenerator(HashGridStorage& s) : storage_(s) {}
  // regular sparse grid of level n (HashGenerator.hpp:282-342, T = 0)
  void regular(size_t level);
  // surplus refinement (HashRefinement.cpp:124-141); new points are appended
  void refine(RefinementFunctor& functor, std::vector<size_t>* addedPoints = nullptr);
  size_t getNumberOfRefinablePoints();

 private:
  HashGridStorage& storage_;
};

class Grid {
 public:
  static Grid* createLinearGrid(size_t dim) { return new Grid(GridType::Linear, dim); }
  static Grid* createModLinearGrid(size_t dim) { return new Grid(GridType::ModLinear, dim); }
  Grid(GridType t, size_t dim) : type_(t), storage_(dim), generator_(storage_) {}
  virtual ~Grid() {}
  GridType getType() const { return type_; }
  HashGridStorage& getStorage() { return storage_; }
  GridGenerator& getGenerator() { return generator_; }
  size_t getSize() const { return storage_.getSize(); }
  size_t getDimension() const { return storage_.getDimension(); }
  // "linear" / "modlinear" + storage text (Grid.cpp:941-958)
  std::string serialize() const;
  static Grid* unserialize(const std::string& text);

 private:
  GridType type_;
  HashGridStorage storage_;
  GridGenerator generator_;
};

// Hierarchisation of nodal values by 1-d sweeps in every dimension
// (OperationHierarchisationLinear / ...ModLinear with the sweep functors of
// base/src/sgpp/base/operation/hash/common/algorithm_sweep/Hierarchisation{Linear,ModLinear}.cpp).
void doHierarchisation(GridType type, HashGridStorage& storage, DataVector& node_values);

// Hierarchisation that keeps its per-sweep intermediates so that points
// appended by refinement are hierarchised without re-sweeping the grid.
// On a grid whose points all have their hierarchical ancestors (what regular
// generation and refinement produce) the surplus of a point depends only on
// the values at its ancestors, and this class performs the same FP64
// operations on them as the recursive sweep, so the surpluses are identical
// bit for bit to a full doHierarchisation.
class IncrementalHierarchisation {
 public:
  IncrementalHierarchisation(GridType type, HashGridStorage& storage)
      : type_(type), storage_(storage) {}
  // node_values holds f at every grid point (size == grid size); points
  // [processed(), size) are hierarchised, alpha is resized and filled.
  void update(const DataVector& node_values, DataVector& alpha);
  size_t processed() const { return done_; }

 private:
  GridType type_;
  HashGridStorage& storage_;
  size_t done_ = 0;
  // stage values v^(k)(p) for k = 0..d, (d+1) per point
  std::vector<double> stages_;
};

}  // namespace base
}  // namespace sgpp
