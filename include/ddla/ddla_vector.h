// ddla::vector<T>: contiguous, uninitialized storage in the handle's memory
// (device memory for the GPU handle), the interface of the reference's
// include/ddla/ddla_vector.h:27-110.  Growth allocates anew, copies the
// existing elements on the handle's stream and releases the old block;
// shrinking and clear() keep the allocation; reset() releases it (errors
// throw); the destructor releases it without throwing.  Move-only; the handle
// must outlive the vector.  There is no host element access: pass data() to
// the library or copy with ddlaMemcpy and synchronize.
#pragma once

#include "ddla.h"

#include <cstddef>
#include <limits>
#include <stdexcept>
#include <type_traits>
#include <utility>

namespace ddla {

template <typename T>
class vector {
    static_assert(std::is_trivially_copyable<T>::value, "ddla::vector holds trivially copyable elements");
    static_assert(!std::is_const<T>::value && !std::is_volatile<T>::value, "ddla::vector holds unqualified types");
    static_assert(alignof(T) <= alignof(std::max_align_t), "ddla::vector does not over-align");

public:
    explicit vector(DdlaHandle_t handle, std::size_t count = 0) : handle_(handle)
    {
        if (handle_ == nullptr) throw std::invalid_argument("ddla::vector: null handle");
        resize(count);
    }
    ~vector() noexcept { release_quietly(); }

    vector(const vector&) = delete;
    vector& operator=(const vector&) = delete;
    vector(vector&& other) noexcept { swap(other); }
    vector& operator=(vector&& other) noexcept
    {
        if (this != &other) {
            vector taken(std::move(other));
            swap(taken);
        }
        return *this;
    }

    DdlaHandle_t handle() const noexcept { return handle_; }
    T* data() noexcept { return data_; }
    const T* data() const noexcept { return data_; }
    std::size_t size() const noexcept { return size_; }
    std::size_t capacity() const noexcept { return capacity_; }
    bool empty() const noexcept { return size_ == 0; }
    std::size_t max_size() const noexcept { return std::numeric_limits<std::size_t>::max() / sizeof(T); }

    void reserve(std::size_t count)
    {
        if (count > max_size()) throw std::length_error("ddla::vector: allocation size overflow");
        if (count <= capacity_) return;
        void* fresh = nullptr;
        DDLA_CHECK(ddlaMalloc(handle_, &fresh, count * sizeof(T)));
        if (size_ != 0) {
            const ddlaStatus_t st = ddlaMemcpy(handle_, fresh, data_, size_ * sizeof(T));
            if (st != ddlaStatus_t::DDLA_STATUS_SUCCESS) {
                (void)ddlaFree(handle_, fresh);
                DDLA_CHECK(st);
            }
        }
        if (data_ != nullptr) DDLA_CHECK(ddlaFree(handle_, data_));
        data_ = static_cast<T*>(fresh);
        capacity_ = count;
    }
    void resize(std::size_t count)
    {
        reserve(count);
        size_ = count;
    }
    void clear() noexcept { size_ = 0; }
    void reset()
    {
        if (data_ != nullptr) DDLA_CHECK(ddlaFree(handle_, data_));
        data_ = nullptr;
        size_ = capacity_ = 0;
    }
    void swap(vector& other) noexcept
    {
        std::swap(handle_, other.handle_);
        std::swap(data_, other.data_);
        std::swap(size_, other.size_);
        std::swap(capacity_, other.capacity_);
    }

private:
    void release_quietly() noexcept
    {
        if (data_ != nullptr) (void)ddlaFree(handle_, data_);
        data_ = nullptr;
    }

    DdlaHandle_t handle_ = nullptr;
    T* data_ = nullptr;
    std::size_t size_ = 0;
    std::size_t capacity_ = 0;
};

} // namespace ddla
