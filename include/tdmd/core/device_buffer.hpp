// device_buffer.hpp -- DeviceBuffer<T> (src/core/device_buffer.cuh:29-116):
// move-only owning device allocation with synchronous host copies, over the
// C-ABI runtime calls. zero() is a kernel fill (no hipMemset).
#pragma once

#include <cstddef>
#include <utility>

#include "error.hpp"

namespace tdmd {

template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(std::size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), size_(o.size_) {
    o.ptr_ = nullptr;
    o.size_ = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = std::exchange(o.ptr_, nullptr);
      size_ = std::exchange(o.size_, 0);
    }
    return *this;
  }

  // Reallocates (contents not kept) when the size changes.
  void resize(std::size_t n) {
    if (n == size_) return;
    release();
    if (n == 0) return;
    void* p = nullptr;
    TDMD_HIP_CALL(tdmd_malloc(&p, n * sizeof(T)));
    ptr_ = static_cast<T*>(p);
    size_ = n;
  }
  void copy_from_host(const T* src, std::size_t n) {
    TDMD_HIP_CALL(tdmd_memcpy_htod(ptr_, src, n * sizeof(T), nullptr));
  }
  void copy_to_host(T* dst, std::size_t n) const {
    TDMD_HIP_CALL(tdmd_memcpy_dtoh(dst, ptr_, n * sizeof(T), nullptr));
  }
  void zero() {
    static_assert(sizeof(T) % 4 == 0, "zero() fills 32-bit words");
    TDMD_HIP_CALL(tdmd_fill_u32(ptr_, 0u, size_ * sizeof(T) / 4, nullptr));
    TDMD_HIP_CALL(tdmd_device_synchronize());
  }

  T* data() noexcept { return ptr_; }
  const T* data() const noexcept { return ptr_; }
  std::size_t size() const noexcept { return size_; }

 private:
  void release() noexcept {
    if (ptr_) (void)tdmd_free(ptr_);
    ptr_ = nullptr;
    size_ = 0;
  }
  T* ptr_ = nullptr;
  std::size_t size_ = 0;
};

}  // namespace tdmd
