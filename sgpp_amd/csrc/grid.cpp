// Grid storage, regular generation, surplus refinement and hierarchisation.
// The operation sequences follow SG++ so that sequence numbers and surpluses
// agree bit for bit; each function cites the reference routine it follows.

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <limits>
#include <sstream>

#include "sgpp/base/grid/Grid.hpp"

namespace sgpp {
namespace base {

// ------------------------------------------------------------------ storage

void HashGridStorage::clear() {
  map_.clear();
  for (HashGridPoint* p : list_) delete p;
  list_.clear();
}

// HashGridStorage.cpp:288-292: append a copy, map it to its position
size_t HashGridStorage::insert(const HashGridPoint& p) {
  if (p.getDimension() != dim_) throw generation_exception("point dimension mismatch");
  HashGridPoint* ins = new HashGridPoint(p);
  list_.push_back(ins);
  return (map_[ins] = list_.size() - 1);
}

// HashGridStorage.cpp:325-336: erase the old key, store a copy at pos
void HashGridStorage::update(const HashGridPoint& p, size_t pos) {
  if (pos < list_.size()) {
    HashGridPoint* del = list_[pos];
    map_.erase(del);
    delete del;
    HashGridPoint* ins = new HashGridPoint(p);
    list_[pos] = ins;
    map_[ins] = pos;
  }
}

void HashGridStorage::getLevelIndexArrays(std::vector<uint32_t>& level,
                                          std::vector<uint32_t>& index) const {
  level.resize(list_.size() * dim_);
  index.resize(list_.size() * dim_);
  for (size_t j = 0; j < list_.size(); ++j)
    for (size_t d = 0; d < dim_; ++d) list_[j]->get(d, level[j * dim_ + d], index[j * dim_ + d]);
}

void HashGridStorage::getLevelIndexArraysForEval(std::vector<double>& level,
                                                 std::vector<double>& index) const {
  level.resize(list_.size() * dim_);
  index.resize(list_.size() * dim_);
  for (size_t j = 0; j < list_.size(); ++j)
    for (size_t d = 0; d < dim_; ++d) {
      level_t l;
      index_t i;
      list_[j]->get(d, l, i);
      level[j * dim_ + d] = static_cast<double>(1u << l);
      index[j * dim_ + d] = static_cast<double>(i);
    }
}

void HashGridStorage::getLevelIndexMaskArraysForModEval(std::vector<double>& level,
                                                        std::vector<double>& index,
                                                        std::vector<uint64_t>& mask,
                                                        std::vector<double>& offset) const {
  const size_t n = list_.size() * dim_;
  level.resize(n);
  index.resize(n);
  mask.resize(n);
  offset.resize(n);
  for (size_t j = 0; j < list_.size(); ++j)
    for (size_t d = 0; d < dim_; ++d) {
      level_t l;
      index_t i;
      list_[j]->get(d, l, i);
      const size_t e = j * dim_ + d;
      if (l == 1) {
        level[e] = 0.0;
        index[e] = 0.0;
        mask[e] = 0;
        offset[e] = 1.0;
      } else if (i == 1) {
        level[e] = -static_cast<double>(1u << l);
        index[e] = 0.0;
        mask[e] = 0;
        offset[e] = 2.0;
      } else if (i == (1u << l) - 1) {
        level[e] = static_cast<double>(1u << l);
        index[e] = static_cast<double>(i);
        mask[e] = 0;
        offset[e] = 1.0;
      } else {
        level[e] = static_cast<double>(1u << l);
        index[e] = static_cast<double>(i);
        mask[e] = 0x8000000000000000ull;
        offset[e] = 1.0;
      }
    }
}

void HashGridStorage::getCoordinateArrays(std::vector<double>& coords) const {
  coords.resize(list_.size() * dim_);
  for (size_t j = 0; j < list_.size(); ++j)
    for (size_t d = 0; d < dim_; ++d) coords[j * dim_ + d] = list_[j]->getStandardCoordinate(d);
}

size_t HashGridStorage::getMaxLevel() const {
  level_t m = 0;
  for (const HashGridPoint* p : list_) m = std::max(m, p->getLevelMax());
  return m;
}

// Text layout of HashGridStorage::serialize (HashGridStorage.cpp:217-243) with
// the unit bounding box and HashGridPoint::serialize (HashGridPoint.cpp:108-123),
// format version 9.
std::string HashGridStorage::serialize() const {
  std::ostringstream os;
  os << 9 << " " << dim_ << " " << list_.size() << std::endl;
  os << std::scientific << 0 << std::endl;
  // BoundingBox::serialize, version >= 9: one "left right bLeft bRight" per dim
  for (size_t d = 0; d < dim_; ++d) os << 0.0 << " " << 1.0 << " " << 0 << " " << 0 << " ";
  os << std::endl;
  for (const HashGridPoint* p : list_) {
    os << dim_ << std::endl;
    for (size_t d = 0; d < dim_; ++d) os << p->getLevel(d) << " " << p->getIndex(d) << " ";
    os << std::endl << (p->isLeaf() ? 1 : 0) << std::endl;
  }
  return os.str();
}

// HashGridStorage::parseGridDescription (HashGridStorage.cpp:605-684) and
// HashGridPoint(std::istream&, version) (HashGridPoint.cpp:52-89), versions 1
// to 9 (SerializationVersion.hpp:11-24): versions 1-2 carry no bounding box,
// 3-4 one "left right bLeft bRight" per dimension, 5 and later a stretching
// flag first (only 0, a bounding box, is supported); points carry a leaf flag
// from version 2 on (not in 4) and a point-distribution field in 6-8.
// Versions 1 and 4 recompute the leaf flags. Only the unit box is accepted.
void HashGridStorage::unserialize(const std::string& text) {
  std::istringstream is(text);
  int version = 0;
  size_t dim = 0, n = 0;
  is >> version >> dim >> n;
  if (!is || version < 1 || (version > 9 && version != 4))
    throw generation_exception("unsupported grid serialization version");
  if (dim != dim_) throw generation_exception("grid dimension mismatch in serialization");
  if (version >= 5) {
    int stretch = 0;
    is >> stretch;
    if (stretch != 0) throw generation_exception("stretched grids are not supported");
  }
  if (version >= 3) {
    for (size_t d = 0; d < dim_; ++d) {
      double lft, rgt;
      int bl, br;
      is >> lft >> rgt >> bl >> br;
      if (lft != 0.0 || rgt != 1.0)
        throw generation_exception("only the unit bounding box is supported");
    }
  }
  clear();
  HashGridPoint p(dim_);
  for (size_t j = 0; j < n; ++j) {
    size_t pd = 0;
    is >> pd;
    if (pd != dim_) throw generation_exception("grid point dimension mismatch");
    for (size_t d = 0; d < dim_; ++d) {
      level_t l;
      index_t i;
      is >> l >> i;
      p.push(d, l, i);
    }
    int leaf = 0;
    if (version >= 2 && version != 4) is >> leaf;
    if (version >= 6 && version <= 8) {
      size_t distribution;
      is >> distribution;
    }
    p.setLeaf(leaf != 0);
    if (!is) throw generation_exception("truncated grid serialization");
    insert(p);
  }
  if (version == 1 || version == 4) recalcLeafProperty();
}

// HashGridStorage::recalcLeafProperty (HashGridStorage.cpp:358-397): a point
// is a leaf when neither child in any dimension (for level 0: the level-1
// point) is in the grid.
void HashGridStorage::recalcLeafProperty() {
  for (HashGridPoint* stored : list_) {
    HashGridPoint q(*stored);
    bool leaf = true;
    for (size_t d = 0; d < dim_; ++d) {
      level_t l;
      index_t i;
      q.get(d, l, i);
      if (l > 0) {
        q.set(d, l + 1, 2 * i - 1);
        leaf = leaf && !isContaining(q);
        q.set(d, l + 1, 2 * i + 1);
        leaf = leaf && !isContaining(q);
      } else {
        q.set(d, 1, 1);
        leaf = leaf && !isContaining(q);
      }
      q.set(d, l, i);
    }
    stored->setLeaf(leaf);
  }
}

std::string Grid::serialize() const {
  std::ostringstream os;
  os << (type_ == GridType::Linear ? "linear" : "modlinear") << std::endl;
  os << storage_.serialize();
  return os.str();
}

Grid* Grid::unserialize(const std::string& text) {
  std::istringstream is(text);
  std::string name;
  is >> name;
  GridType t;
  if (name == "linear")
    t = GridType::Linear;
  else if (name == "modlinear")
    t = GridType::ModLinear;
  else
    throw factory_exception("grid type '" + name + "' is not supported");
  std::string rest((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
  std::istringstream hs(rest);
  int version = 0;
  size_t dim = 0;
  hs >> version >> dim;
  Grid* g = new Grid(t, dim);
  try {
    g->getStorage().unserialize(rest);
  } catch (...) {
    delete g;
    throw;
  }
  return g;
}

// -------------------------------------------------------------- generation

// HashGenerator::regular_iter with T = 0 (HashGenerator.hpp:282-342)
void GridGenerator::regular(size_t n) {
  HashGridStorage& st = storage_;
  const size_t dim = st.getDimension();
  if (dim == 0) return;
  if (n < 1 || n > 30) throw generation_exception("regular grid level must be in [1, 30]");
  HashGridPoint idx1(dim);
  for (size_t d = 0; d < dim; ++d) idx1.push(d, 1, 1, false);
  for (level_t l = 1; l <= n; ++l)
    for (index_t i = 1; i < (index_t(1) << l); i += 2) {
      idx1.push(0, l, i, l == n);
      st.insert(idx1);
    }
  const level_t top = static_cast<level_t>(n + dim - 1);
  for (size_t d = 1; d < dim; ++d) {
    const size_t gsize = st.getSize();
    for (size_t g = 0; g < gsize; ++g) {
      bool first = true;
      HashGridPoint idx(st.getPoint(g));
      const level_t lsum = idx.getLevelSum() - 1;
      const level_t lmax = idx.getLevelMax();
      for (level_t l = 1; (l + lsum <= top) && (std::max(l, lmax) <= n); ++l)
        for (index_t i = 1; i < (index_t(1) << l); i += 2) {
          idx.push(d, l, i, (l + lsum) == top);
          if (!first) {
            st.insert(idx);
          } else {
            st.update(idx, g);
            first = false;
          }
        }
    }
  }
}

double SurplusRefinementFunctor::operator()(HashGridStorage&, size_t seq) const {
  return std::fabs(alpha_[seq]);
}

namespace {

typedef std::pair<size_t, double> ref_pair;  // (sequence number, indicator)

// AbstractRefinement::compare_pairs (AbstractRefinement.hpp:120-123)
bool compare_pairs(const ref_pair& a, const ref_pair& b) { return a.second > b.second; }

// AbstractRefinement::createGridpoint1D + createGridpointSubroutine +
// HashRefinement::createGridpoint: ancestors first, depth first by dimension
void create_gridpoint(HashGridStorage& st, HashGridPoint& point);

void create_subroutine(HashGridStorage& st, HashGridPoint& point) {
  if (!st.isContaining(point)) {
    const bool save = point.isLeaf();
    point.setLeaf(false);
    create_gridpoint(st, point);
    point.setLeaf(save);
  } else {
    st.getPoint(st.find(&point)->second).setLeaf(false);
  }
}

void create_gridpoint(HashGridStorage& st, HashGridPoint& point) {
  for (size_t d = 0; d < st.getDimension(); ++d) {
    level_t l;
    index_t i;
    point.get(d, l, i);
    if (l > 1) {
      if (((i + 1) / 2) % 2 == 1)
        point.set(d, l - 1, (i + 1) / 2);
      else
        point.set(d, l - 1, (i - 1) / 2);
      create_subroutine(st, point);
      point.set(d, l, i);
    }
  }
  st.insert(point);
}

// HashRefinement::refineGridpoint1D (HashRefinement.cpp:203-221)
void refine_1d(HashGridStorage& st, HashGridPoint& point, size_t d) {
  level_t l;
  index_t i;
  point.get(d, l, i);
  point.set(d, l + 1, 2 * i - 1);
  if (!st.isContaining(point)) {
    point.setLeaf(true);
    create_gridpoint(st, point);
  }
  point.set(d, l + 1, 2 * i + 1);
  if (!st.isContaining(point)) {
    point.setLeaf(true);
    create_gridpoint(st, point);
  }
  point.set(d, l, i);
}

// first missing child in any dimension (HashRefinement.cpp:60-93)
bool has_missing_child(HashGridStorage& st, HashGridPoint& point) {
  for (size_t d = 0; d < st.getDimension(); ++d) {
    level_t l;
    index_t i;
    point.get(d, l, i);
    point.set(d, l + 1, 2 * i - 1);
    if (!st.isContaining(point)) return true;
    point.set(d, l + 1, 2 * i + 1);
    if (!st.isContaining(point)) return true;
    point.set(d, l, i);
  }
  return false;
}

}  // namespace

// HashRefinement::free_refine (HashRefinement.cpp:124-141) with
// collectRefinablePoints (:42-95), addElementToCollection (:19-40) and
// refineGridpointsCollection (:111-122)
void GridGenerator::refine(RefinementFunctor& functor, std::vector<size_t>* addedPoints) {
  HashGridStorage& st = storage_;
  if (st.getSize() == 0) throw generation_exception("storage empty");
  const size_t before = st.getSize();
  const size_t num = functor.getRefinementsNum();
  std::vector<ref_pair> collection;
  HashGridPoint point(st.getDimension());
  for (auto it = st.begin(); it != st.end(); ++it) {
    if (it->first->childrenComplete()) continue;
    point = *(it->first);
    if (!has_missing_child(st, point)) {
      it->first->setChildrenComplete();
      continue;
    }
    collection.push_back(ref_pair(it->second, functor(st, it->second)));
    std::push_heap(collection.begin(), collection.end(), compare_pairs);
    if (collection.size() > num) {
      std::pop_heap(collection.begin(), collection.end(), compare_pairs);
      collection.pop_back();
    }
  }
  const double thr = functor.getRefinementThreshold();
  // children of a level-31 point would need level 32 and index 2^32 - 1 or
  // more in some later step; the reference's 32-bit index silently wraps
  // there, so the limit is checked before the grid is touched
  for (const ref_pair& pr : collection) {
    if (pr.second < thr) continue;
    const HashGridPoint& p = st.getPoint(pr.first);
    for (size_t d = 0; d < st.getDimension(); ++d)
      if (p.getLevel(d) >= kMaxRefineLevel)
        throw generation_exception("refinement would create a point above level " +
                                   std::to_string(kMaxRefineLevel));
  }
  for (const ref_pair& pr : collection) {
    if (pr.second >= thr) {
      // HashRefinement::refineGridpoint (HashRefinement.cpp:228-236)
      HashGridPoint p(st[pr.first]);
      st[pr.first].setLeaf(false);
      for (size_t d = 0; d < st.getDimension(); ++d) refine_1d(st, p, d);
    }
  }
  if (addedPoints)
    for (size_t i = before; i < st.getSize(); ++i) addedPoints->push_back(i);
}

size_t GridGenerator::getNumberOfRefinablePoints() {
  size_t c = 0;
  HashGridPoint point(storage_.getDimension());
  for (auto it = storage_.begin(); it != storage_.end(); ++it) {
    if (it->first->childrenComplete()) continue;
    point = *(it->first);
    if (has_missing_child(storage_, point))
      ++c;
    else
      it->first->setChildrenComplete();
  }
  return c;
}

// --------------------------------------------------------- hierarchisation

// Stage k+1 of point p: v_{k+1}(p) = v_k(p) - (fl + fr) / 2, where (fl, fr)
// are the boundary values the recursive sweep in dimension k hands down the
// path from level 1 to p (HierarchisationLinear.cpp rec, and the ModLinear
// extrapolation of HierarchisationModLinear.cpp rec for l == 1, i == 1 and
// i == 2^l - 1).
void IncrementalHierarchisation::update(const DataVector& node_values, DataVector& alpha) {
  HashGridStorage& st = storage_;
  const size_t n = st.getSize();
  const size_t dim = st.getDimension();
  if (node_values.size() != n) throw data_exception("node_values size != grid size");
  if (done_ > n) throw generation_exception("grid shrank since the last hierarchisation");
  stages_.resize(n * (dim + 1));
  // process new points by level sum, ties by sequence number
  std::vector<size_t> order;
  order.reserve(n - done_);
  for (size_t s = done_; s < n; ++s) order.push_back(s);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return st.getPoint(a).getLevelSum() < st.getPoint(b).getLevelSum();
  });
  HashGridPoint probe(dim);
  for (size_t s : order) {
    double* v = &stages_[s * (dim + 1)];
    v[0] = node_values[s];
    const HashGridPoint& p = st.getPoint(s);
    for (size_t k = 0; k < dim; ++k) {
      level_t lp;
      index_t ip;
      p.get(k, lp, ip);
      double fl = 0.0, fr = 0.0;
      probe = p;
      for (level_t la = 1; la < lp; ++la) {
        const index_t ia = (ip >> (lp - la)) | 1u;
        probe.set(k, la, ia);
        const size_t sa = st.getSequenceNumber(probe);
        if (st.isInvalidSequenceNumber(sa))
          throw generation_exception("grid is missing a hierarchical ancestor");
        const double fm = stages_[sa * (dim + 1) + k];
        double flt = fl, frt = fr;
        if (type_ == GridType::ModLinear) {
          if (la == 1) {
            flt = fm;
            frt = fm;
          } else if (ia == 1) {
            const double ft = fr - fm;
            flt = fm - ft;
          } else if (ia == (index_t(1) << la) - 1) {
            const double ft = fl - fm;
            frt = fm - ft;
          }
        }
        const index_t child = (ip >> (lp - la - 1)) | 1u;
        if (child == 2 * ia - 1) {
          fl = flt;
          fr = fm;
        } else {
          fl = fm;
          fr = frt;
        }
      }
      v[k + 1] = v[k] - ((fl + fr) / 2.0);
    }
  }
  done_ = n;
  alpha.resize(n);
  for (size_t s = 0; s < n; ++s) alpha[s] = stages_[s * (dim + 1) + dim];
}

void doHierarchisation(GridType type, HashGridStorage& storage, DataVector& node_values) {
  IncrementalHierarchisation h(type, storage);
  DataVector alpha;
  h.update(node_values, alpha);
  node_values = alpha;
}

}  // namespace base
}  // namespace sgpp
