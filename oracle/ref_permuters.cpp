// ref_permuters.cpp -- exports the reference's OWN permuters for checking.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`)
// against the unmodified Jet headers where they lie
// (include/jet/permute/{Permuter,Default,QFlex}.hpp and their only
// dependencies Abort.hpp / Utilities.hpp, all STL-only).  The result goes to
// oracle/_ref/libjet_ref.so, which tests use to pin the oracle's QFlex
// restatement and the GPU permutation against the real implementation.
//
// This file is a C-ABI wrapper written for this project: it instantiates the
// template configurations the reference's own tests sweep
// (test/Test_Permuter.cpp:182-190, 458-465, 492-496) and turns Jet::Exception
// into an error string.
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "jet/permute/PermuterIncludes.hpp"

namespace {

using Data = std::complex<float>;

template <class Backend>
int run(int rank, const int64_t *shape, const char *const *old_labels,
        const char *const *new_labels, const Data *in, size_t n_in, Data *out,
        char *err, size_t errlen)
{
    std::vector<size_t> sh(shape, shape + rank);
    std::vector<std::string> o(old_labels, old_labels + rank);
    std::vector<std::string> nw;
    for (int i = 0; new_labels[i] != nullptr; i++) nw.emplace_back(new_labels[i]);
    std::vector<Data> data(in, in + n_in);
    try {
        Jet::Permuter<Backend> p;
        auto r = p.Transpose(data, sh, o, nw);
        std::memcpy(out, r.data(), sizeof(Data) * r.size());
        return 0;
    }
    catch (const Jet::Exception &e) {
        std::strncpy(err, e.what(), errlen - 1);
        err[errlen - 1] = 0;
        return 1;
    }
}

} // namespace

extern "C" {

// backend: 0 = DefaultPermuter<blocksize>, 1 = QFlexPermuter<blocksize,
// min_dims>.  new_labels is NULL-terminated so that its length may differ
// from rank (the reference's validation then reports the mismatch).
// Returns 0 on success, 1 on Jet::Exception (message in err), 2 when the
// template configuration is not instantiated here.
int jr_permute(int backend, size_t blocksize, size_t min_dims, int rank,
               const int64_t *shape, const char *const *old_labels,
               const char *const *new_labels, const void *in, size_t n_in,
               void *out, char *err, size_t errlen)
{
    auto *i = static_cast<const Data *>(in);
    auto *o = static_cast<Data *>(out);
#define JR_CALL(...) return run<__VA_ARGS__>(rank, shape, old_labels, new_labels, i, n_in, o, err, errlen)
    if (backend == 0) {
        switch (blocksize) {
        case 64: JR_CALL(Jet::DefaultPermuter<64>);
        case 128: JR_CALL(Jet::DefaultPermuter<128>);
        case 256: JR_CALL(Jet::DefaultPermuter<256>);
        case 512: JR_CALL(Jet::DefaultPermuter<512>);
        case 1024: JR_CALL(Jet::DefaultPermuter<1024>);
        case 2048: JR_CALL(Jet::DefaultPermuter<2048>);
        default: return 2;
        }
    }
    if (min_dims == 32) {
        switch (blocksize) {
        case 64: JR_CALL(Jet::QFlexPermuter<64, 32>);
        case 128: JR_CALL(Jet::QFlexPermuter<128, 32>);
        case 256: JR_CALL(Jet::QFlexPermuter<256, 32>);
        case 512: JR_CALL(Jet::QFlexPermuter<512, 32>);
        case 1024: JR_CALL(Jet::QFlexPermuter<1024, 32>);
        case 2048: JR_CALL(Jet::QFlexPermuter<2048, 32>);
        default: return 2;
        }
    }
    if (blocksize == 1024) {
        switch (min_dims) {
        case 2: JR_CALL(Jet::QFlexPermuter<1024, 2>);
        case 64: JR_CALL(Jet::QFlexPermuter<1024, 64>);
        case 128: JR_CALL(Jet::QFlexPermuter<1024, 128>);
        case 256: JR_CALL(Jet::QFlexPermuter<1024, 256>);
        case 512: JR_CALL(Jet::QFlexPermuter<1024, 512>);
        default: return 2;
        }
    }
    if (blocksize == 2048 && min_dims == 1024) JR_CALL(Jet::QFlexPermuter<2048, 1024>);
#undef JR_CALL
    return 2;
}

} // extern "C"
