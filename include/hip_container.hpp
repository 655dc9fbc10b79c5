// hip_container<T>: a host std::vector mirrored by a device array, the
// HIP-native stand-in for the reference's cuda_container<T>
// (include/cuda_container.hcu:18-88, src/cuda_container.cu).  Same member
// names: h_array(), d_array().data(), size(), resize(), set(),
// transfer_to_device(), transfer_to_host(), operator[], at().
//
// Differences, on purpose: resize() zeroes the new device memory, and a
// container can be a non-owning view of device memory owned elsewhere (the
// engine's accumulators returned by glst_force<CT>::fx()).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

template <typename T> class device_span {
 public:
  T *data() const { return ptr_; }
  std::size_t size() const { return n_; }

 private:
  template <typename U> friend class hip_container;
  T *ptr_ = nullptr;
  std::size_t n_ = 0;
};

template <typename T> class hip_container {
 public:
  hip_container() = default;
  explicit hip_container(std::size_t count) { resize(count); }
  explicit hip_container(const std::vector<T> &v) { set(v); }
  hip_container(const hip_container &o) { *this = o; }
  hip_container(hip_container &&o) noexcept { swap(o); }
  hip_container &operator=(const hip_container &o) {
    if (this != &o) {
      h_ = o.h_;
      resize(o.size());
      if (o.d_.n_)
        check(hipMemcpy(d_.ptr_, o.d_.ptr_, o.d_.n_ * sizeof(T),
                        hipMemcpyDeviceToDevice));
    }
    return *this;
  }
  hip_container &operator=(hip_container &&o) noexcept {
    swap(o);
    return *this;
  }
  hip_container &operator=(const std::vector<T> &v) {
    set(v);
    return *this;
  }
  ~hip_container() { release(); }

  // Non-owning view of n elements of device memory at dptr.
  static hip_container view(T *dptr, std::size_t n) {
    hip_container c;
    c.h_.resize(n);
    c.d_.ptr_ = dptr;
    c.d_.n_ = n;
    c.owning_ = false;
    return c;
  }

  // element access (host side, as in the reference)
  const T &at(std::size_t i) const { return h_.at(i); }
  T &at(std::size_t i) { return h_.at(i); }
  const T &operator[](std::size_t i) const { return h_[i]; }
  T &operator[](std::size_t i) { return h_[i]; }
  const std::vector<T> &h_array() const { return h_; }
  std::vector<T> &h_array() { return h_; }
  const device_span<T> &d_array() const { return d_; }
  device_span<T> &d_array() { return d_; }
  std::size_t size() const { return h_.size(); }

  void clear() {
    h_.clear();
    release();
  }

  void resize(std::size_t n) {
    h_.resize(n);
    if (n == d_.n_) return;
    if (!owning_)
      throw std::runtime_error("hip_container: cannot resize a view");
    release();
    if (n) {
      void *p = nullptr;
      check(hipMalloc(&p, n * sizeof(T)));
      d_.ptr_ = static_cast<T *>(p);
      d_.n_ = n;
      // hipMemset may run asynchronously on the null stream; finish it so
      // work queued on other (non-blocking) streams never races it.
      check(hipMemset(p, 0, n * sizeof(T)));
      check(hipDeviceSynchronize());
    }
  }

  void set(const std::vector<T> &v) {
    h_ = v;
    resize(v.size());
    transfer_to_device();
  }
  void set(const T &value) {
    h_.assign(h_.size(), value);
    transfer_to_device();
  }

  void transfer_to_device() {
    if (!h_.empty())
      check(hipMemcpy(d_.ptr_, h_.data(), h_.size() * sizeof(T),
                      hipMemcpyHostToDevice));
    check(hipDeviceSynchronize());
  }
  void transfer_to_host() {
    if (!h_.empty())
      check(hipMemcpy(h_.data(), d_.ptr_, h_.size() * sizeof(T),
                      hipMemcpyDeviceToHost));
    check(hipDeviceSynchronize());
  }

 private:
  static void check(hipError_t e) {
    if (e != hipSuccess)
      throw std::runtime_error(std::string("HIP Runtime Error: ") +
                               hipGetErrorString(e));
  }
  void release() {
    if (owning_ && d_.ptr_) (void)hipFree(d_.ptr_);
    d_.ptr_ = nullptr;
    d_.n_ = 0;
  }
  void swap(hip_container &o) noexcept {
    std::swap(h_, o.h_);
    std::swap(d_, o.d_);
    std::swap(owning_, o.owning_);
  }

  std::vector<T> h_;
  device_span<T> d_;
  bool owning_ = true;
};
