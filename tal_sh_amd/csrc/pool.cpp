#include "pool.hpp"

#include "tensor_algebra.h"

namespace tal {

void Arena::init(void *base, size_t size, size_t align) {
  base_ = static_cast<char *>(base);
  align_ = align ? align : 256;
  size_ = (size / align_) * align_;
  reset();
}

void Arena::reset() {
  by_off_.clear();
  by_len_.clear();
  slots_.clear();
  free_slots_.clear();
  free_total_ = 0;
  live_ = 0;
  if (size_ > 0) insert_free(0, size_);
}

void Arena::insert_free(size_t off, size_t len) {
  by_off_[off] = len;
  by_len_.emplace(len, off);
  free_total_ += len;
}

void Arena::erase_free(std::map<size_t, size_t>::iterator it) {
  auto range = by_len_.equal_range(it->second);
  for (auto j = range.first; j != range.second; ++j)
    if (j->second == it->first) {
      by_len_.erase(j);
      break;
    }
  free_total_ -= it->second;
  by_off_.erase(it);
}

size_t Arena::largest_free() const { return by_len_.empty() ? 0 : by_len_.rbegin()->first; }

size_t Arena::length(int handle) const {
  if (handle < 0 || handle >= static_cast<int>(slots_.size()) || !slots_[handle].used) return 0;
  return slots_[handle].len;
}

bool Arena::owns(const void *p) const {
  const char *c = static_cast<const char *>(p);
  return base_ && c >= base_ && c < base_ + size_;
}

int Arena::alloc(size_t bytes, void **ptr, int *handle) {
  size_t len = ((bytes ? bytes : 1) + align_ - 1) / align_ * align_;
  if (len > size_) return DEVICE_UNABLE;
  auto it = by_len_.lower_bound(len);  // best fit: smallest free range that fits
  if (it == by_len_.end()) return TRY_LATER;
  const size_t off = it->second, flen = it->first;
  erase_free(by_off_.find(off));
  if (flen > len) insert_free(off + len, flen - len);
  int h;
  if (!free_slots_.empty()) {
    h = free_slots_.back();
    free_slots_.pop_back();
  } else {
    h = static_cast<int>(slots_.size());
    slots_.emplace_back();
  }
  slots_[h].off = off;
  slots_[h].len = len;
  slots_[h].used = true;
  ++live_;
  *ptr = base_ + off;
  *handle = h;
  return 0;
}

int Arena::release(int handle) {
  if (handle < 0 || handle >= static_cast<int>(slots_.size()) || !slots_[handle].used) return -1;
  size_t off = slots_[handle].off, len = slots_[handle].len;
  slots_[handle].used = false;
  free_slots_.push_back(handle);
  --live_;
  // merge with the free neighbours
  auto next = by_off_.lower_bound(off);
  if (next != by_off_.end() && next->first == off + len) {
    len += next->second;
    erase_free(next);
  }
  auto nxt = by_off_.lower_bound(off);
  if (nxt != by_off_.begin()) {
    auto prev = std::prev(nxt);
    if (prev->first + prev->second == off) {
      off = prev->first;
      len += prev->second;
      erase_free(prev);
    }
  }
  insert_free(off, len);
  return 0;
}

}  // namespace tal
