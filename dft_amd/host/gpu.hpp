// gpu.hpp -- thin C++ ownership over the kernel library's C-ABI.  Status
// codes become the reference's exception types: -1 -> std::invalid_argument,
// everything else -> std::runtime_error (the reference reports all errors
// by exception, e.g. EigenvalueSolverMolecule.cpp:1214-1220).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "dft_hip.h"

namespace dftc {

inline void check(dft_ctx* c, int rc, const char* what) {
    if (rc == DFT_OK) return;
    std::string msg = std::string(what) + ": " + (c ? dft_last_error(c) : "no context");
    if (rc == DFT_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

class Ctx {
   public:
    explicit Ctx(int device = 0) {
        const int rc = dft_ctx_create(device, &c_);
        if (rc != DFT_OK) throw std::runtime_error("dft_ctx_create failed: no usable GPU");
    }
    ~Ctx() { dft_ctx_destroy(c_); }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
    dft_ctx* get() const { return c_; }
    operator dft_ctx*() const { return c_; }
    void sync() const { check(c_, dft_sync(c_), "dft_sync"); }

   private:
    dft_ctx* c_ = nullptr;
};

// A device allocation of `bytes` owned by one context.
class Buf {
   public:
    Buf() = default;
    Buf(dft_ctx* c, size_t bytes) : c_(c), bytes_(bytes) { check(c_, dft_alloc(c_, bytes, &p_), "dft_alloc"); }
    ~Buf() { reset(); }
    Buf(Buf&& o) noexcept : c_(o.c_), p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
    Buf& operator=(Buf&& o) noexcept {
        if (this != &o) {
            reset();
            c_ = o.c_; p_ = o.p_; bytes_ = o.bytes_;
            o.p_ = nullptr; o.bytes_ = 0;
        }
        return *this;
    }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    void reset() {
        if (p_) dft_free(c_, p_);
        p_ = nullptr;
        bytes_ = 0;
    }
    void* get() const { return p_; }
    template <typename T>
    T* as() const { return static_cast<T*>(p_); }
    size_t bytes() const { return bytes_; }
    // byte offset helper for column j of a block with `ld` elements of `esz` bytes
    void* col(int64_t j, int64_t ld, size_t esz) const { return static_cast<char*>(p_) + j * ld * (int64_t)esz; }
    void zero() { check(c_, dft_zero(c_, p_, bytes_), "dft_zero"); }
    void upload(const void* h, size_t bytes) { check(c_, dft_upload(c_, p_, h, bytes), "dft_upload"); }
    void download(void* h, size_t bytes) const { check(c_, dft_download(c_, h, p_, bytes), "dft_download"); }

   private:
    dft_ctx* c_ = nullptr;
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

inline size_t esize(int prec) { return prec == DFT_F64 ? 8 : 4; }

}  // namespace dftc
