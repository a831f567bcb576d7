// Bounded rectangle algebra for section coherence.
//
// A Region is a set of pairwise-disjoint half-open boxes in the zero-based physical coordinates of one
// array. The runtime keeps three per registered array: the cells whose values are defined, the cells
// whose host copy is current and the cells whose device copy is current. The limits are those of the
// reference region algebra (compiler/runtime/scoped_regions.hpp:13-15): rank <= 15, at most 128 boxes per
// region; an operation that would exceed them fails and leaves the region unchanged.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace ftc {

constexpr int kMaxRank = 15;
constexpr std::size_t kMaxBoxes = 128;

struct Box {
    int rank = 0;
    std::size_t lo[kMaxRank] = {};
    std::size_t hi[kMaxRank] = {};

    bool empty() const
    {
        for (int d = 0; d < rank; ++d)
            if (lo[d] >= hi[d]) return true;
        return rank == 0;
    }
    std::size_t volume() const
    {
        if (empty()) return 0;
        std::size_t v = 1;
        for (int d = 0; d < rank; ++d) v *= hi[d] - lo[d];
        return v;
    }
    bool intersects(const Box &o) const
    {
        for (int d = 0; d < rank; ++d)
            if (lo[d] >= o.hi[d] || o.lo[d] >= hi[d]) return false;
        return !empty() && !o.empty();
    }
    Box intersect(const Box &o) const
    {
        Box r = *this;
        for (int d = 0; d < rank; ++d) {
            r.lo[d] = lo[d] > o.lo[d] ? lo[d] : o.lo[d];
            r.hi[d] = hi[d] < o.hi[d] ? hi[d] : o.hi[d];
        }
        return r;
    }
    bool contains(const Box &o) const
    {
        for (int d = 0; d < rank; ++d)
            if (o.lo[d] < lo[d] || o.hi[d] > hi[d]) return false;
        return true;
    }
    bool operator==(const Box &o) const
    {
        if (rank != o.rank) return false;
        for (int d = 0; d < rank; ++d)
            if (lo[d] != o.lo[d] || hi[d] != o.hi[d]) return false;
        return true;
    }
};

// a \ b as at most 2*rank disjoint boxes appended to out.
inline void subtract_box(const Box &a, const Box &b, std::vector<Box> &out)
{
    if (!a.intersects(b)) {
        if (!a.empty()) out.push_back(a);
        return;
    }
    Box rest = a;
    for (int d = 0; d < a.rank; ++d) {
        if (rest.lo[d] < b.lo[d]) {
            Box piece = rest;
            piece.hi[d] = b.lo[d];
            out.push_back(piece);
            rest.lo[d] = b.lo[d];
        }
        if (rest.hi[d] > b.hi[d]) {
            Box piece = rest;
            piece.lo[d] = b.hi[d];
            out.push_back(piece);
            rest.hi[d] = b.hi[d];
        }
    }
}

class Region {
  public:
    std::vector<Box> boxes;

    bool empty() const { return boxes.empty(); }
    std::size_t volume() const
    {
        std::size_t v = 0;
        for (const Box &b : boxes) v += b.volume();
        return v;
    }
    // The part of `box` not covered by this region.
    std::vector<Box> missing(const Box &box) const
    {
        std::vector<Box> pieces;
        if (!box.empty()) pieces.push_back(box);
        for (const Box &e : boxes) {
            if (pieces.empty()) break;
            std::vector<Box> next;
            for (const Box &p : pieces) subtract_box(p, e, next);
            pieces.swap(next);
        }
        return pieces;
    }
    bool covers(const Box &box) const { return missing(box).empty(); }
    // The parts of this region inside `box`.
    std::vector<Box> within(const Box &box) const
    {
        std::vector<Box> r;
        for (const Box &e : boxes)
            if (e.intersects(box)) r.push_back(e.intersect(box));
        return r;
    }
    bool intersects(const Box &box) const
    {
        for (const Box &e : boxes)
            if (e.intersects(box)) return true;
        return false;
    }
    // Union with `box`; false (and no change) if the result exceeds the box limit.
    bool add(const Box &box)
    {
        std::vector<Box> fresh = missing(box);
        if (fresh.empty()) return true;
        Region trial = *this;
        trial.boxes.insert(trial.boxes.end(), fresh.begin(), fresh.end());
        trial.coalesce();
        if (trial.boxes.size() > kMaxBoxes) return false;
        boxes.swap(trial.boxes);
        return true;
    }
    bool remove(const Box &box)
    {
        if (!intersects(box)) return true;
        Region trial;
        for (const Box &e : boxes) subtract_box(e, box, trial.boxes);
        trial.coalesce();
        if (trial.boxes.size() > kMaxBoxes) return false;
        boxes.swap(trial.boxes);
        return true;
    }
    void clear() { boxes.clear(); }

    // Merges boxes that share all extents but one and touch along it, until none do.
    void coalesce()
    {
        bool merged = true;
        while (merged) {
            merged = false;
            for (std::size_t a = 0; a < boxes.size() && !merged; ++a)
                for (std::size_t b = a + 1; b < boxes.size() && !merged; ++b) {
                    int differing = -1, count = 0;
                    for (int d = 0; d < boxes[a].rank; ++d)
                        if (boxes[a].lo[d] != boxes[b].lo[d] || boxes[a].hi[d] != boxes[b].hi[d]) {
                            differing = d;
                            ++count;
                        }
                    if (count != 1) continue;
                    Box &x = boxes[a];
                    const Box &y = boxes[b];
                    if (x.hi[differing] == y.lo[differing]) {
                        x.hi[differing] = y.hi[differing];
                    } else if (y.hi[differing] == x.lo[differing]) {
                        x.lo[differing] = y.lo[differing];
                    } else {
                        continue;
                    }
                    boxes.erase(boxes.begin() + static_cast<std::ptrdiff_t>(b));
                    merged = true;
                }
        }
    }
};

}  // namespace ftc
