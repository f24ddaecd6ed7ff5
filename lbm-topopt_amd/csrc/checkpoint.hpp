// Online level-based checkpoint placement with recomputation (Wang, Moin & Iaccarino 2009),
// the schedule of the reference's DynamicCheckpointing + WangCheckpointing
// (src/topOpt/dynamicCheckpointing.cc:9-54, src/topOpt/WangCheckpointing.cc:36-126), and its
// StaticCheckpointing variant for a known horizon (src/topOpt/staticCheckpointing.cc:3-55).
//
// Storage is abstract: the schedule hands out slot ids 0..nslots-1 and the engine maps each
// slot to an HBM buffer. Placing a state in a slot is done by exchanging buffer handles
// (no device copies), so the schedule only has to say which step goes to which slot.
#pragma once
#include <algorithm>
#include <climits>
#include <list>
#include <stdexcept>
#include <vector>

namespace lbmx {

class WangSchedule {
 public:
  struct Entry {
    int step;
    int level;
    int slot;
  };
  // A request for the state at `index`: restart from `from_step` held in `from_slot`,
  // advance, and after producing each step in `stores` put the state into that slot.
  // The state at `index` itself is returned to the caller (not stored).
  struct Plan {
    int from_step = 0;
    int from_slot = -1;
    int index = 0;
    std::vector<std::pair<int, int>> stores;  // (step, slot), increasing step
    int advances() const { return index - from_step; }
    bool stored() const { return from_step == index; }  // no recomputation needed
  };

  explicit WangSchedule(int nslots) : nslots_(nslots) {
    if (nslots < 2) throw std::invalid_argument("dynamic checkpointing needs at least 2 slots");
    reset();
  }

  int nslots() const { return nslots_; }

  void reset() {
    list_.clear();
    free_.clear();
    for (int s = nslots_ - 1; s >= 0; --s) free_.push_back(s);
    have_reader_ = false;
    next_planned_ = 0;
  }

  // StaticCheckpointing: the forward pass keeps exactly the checkpoints that the level-based
  // placement holds after planning the whole horizon 0..last at once (computeNewCheckpoints
  // from the initial state), with their levels; the reverse pass is the same as for the
  // dynamic schedule.
  void plan_static(int last) {
    reset();
    list_.push_front({0, INT_MAX, take()});
    place(0, last);
    planned_.clear();
    for (auto it = list_.rbegin(); it != list_.rend(); ++it) planned_.push_back({it->step, it->level, -1});
    reset();
    static_ = true;
  }
  bool is_static() const { return static_; }

  // DynamicCheckpointing::makeCheckpoint: state `index` always becomes the newest checkpoint.
  // Returns its slot. The slot's previous content (if any) is dropped. A static schedule keeps
  // only its planned steps (StaticCheckpointing::makeCheckpoint) and returns -1 for the others.
  int make_checkpoint(int index) {
    if (static_) {
      if (next_planned_ >= planned_.size() || planned_[next_planned_].step != index) return -1;
      int s = take();
      list_.push_front({index, planned_[next_planned_].level, s});
      ++next_planned_;
      return s;
    }
    if (list_.empty()) {
      int s = take();
      list_.push_front({index, INT_MAX, s});
      reader_ = list_.begin();
      have_reader_ = true;
      return s;
    }
    if (!free_.empty()) {
      int s = take();
      list_.push_front({index, 0, s});
      reader_ = list_.begin();
      have_reader_ = true;
      return s;
    }
    auto movable = std::next(reader_);
    if (movable == list_.end()) throw std::logic_error("checkpoint schedule exhausted");
    if (movable->level < reader_->level) {
      int s = movable->slot;
      list_.push_front({index, 0, s});
      list_.erase(movable);
      return s;
    }
    reader_ = list_.begin();
    int lvl = reader_->level + 1;
    int s = reader_->slot;
    list_.push_front({index, lvl, s});
    list_.erase(reader_);
    reader_ = list_.begin();
    return s;
  }

  // WangCheckpointing::getForwardSolution: plans the delivery of state `index`, placing new
  // checkpoints between the newest stored state and `index` (computeNewCheckpoints). The
  // checkpoint of `index` is consumed; call release() once the state has been used.
  Plan get_forward_solution(int index) {
    Plan p;
    p.index = index;
    if (list_.empty()) throw std::logic_error("no checkpoint stored");
    const Entry start = list_.front();
    if (start.step > index) throw std::logic_error("states must be requested in decreasing order");
    p.from_step = start.step;
    p.from_slot = start.slot;
    if (start.step == index) {
      consumed_slot_ = start.slot;
      list_.pop_front();
      return p;
    }
    place(start.step, index);
    // new entries are those in front of the (former) starting checkpoint, all with step >
    // start.step; the newest one is `index`.
    for (auto& e : list_) {
      if (e.step <= start.step) break;
      if (e.step < index) p.stores.push_back({e.step, e.slot});
    }
    std::reverse(p.stores.begin(), p.stores.end());
    if (list_.front().step != index) throw std::logic_error("schedule lost the requested state");
    consumed_slot_ = list_.front().slot;
    list_.pop_front();
    return p;
  }

  // free the slot consumed by the last get_forward_solution
  void release() {
    if (consumed_slot_ >= 0) free_.push_back(consumed_slot_);
    consumed_slot_ = -1;
  }

  const std::list<Entry>& entries() const { return list_; }
  int free_slots() const { return static_cast<int>(free_.size()); }

 private:
  int take() {
    int s = free_.back();
    free_.pop_back();
    return s;
  }

  // computeNewCheckpoints (src/topOpt/WangCheckpointing.cc:83-126)
  void place(int from, int index) {
    auto reader = list_.begin();
    for (int ii = from; ii < index; ++ii) {
      if (!free_.empty()) {
        list_.push_front({ii + 1, 0, take()});
        reader = list_.begin();
        continue;
      }
      auto movable = std::next(reader);
      if (movable == list_.end()) throw std::logic_error("checkpoint schedule exhausted");
      if (movable->level < reader->level) {
        int s = movable->slot;
        list_.push_front({ii + 1, 0, s});
        list_.erase(movable);
      } else {
        reader = list_.begin();
        int lvl = reader->level + 1;
        int s = reader->slot;
        list_.push_front({ii + 1, lvl, s});
        list_.erase(reader);
        reader = list_.begin();
      }
    }
    // the dynamic placement's reader refers to forward-pass entries only
    have_reader_ = false;
  }

  int nslots_;
  bool static_ = false;
  std::vector<Entry> planned_;  // static schedule: forward checkpoints, increasing step
  size_t next_planned_ = 0;
  std::list<Entry> list_;  // newest first
  std::vector<int> free_;
  std::list<Entry>::iterator reader_;
  bool have_reader_ = false;
  int consumed_slot_ = -1;
};

}  // namespace lbmx
