// hip_tensor_check.cpp -- GPU checks of the header-only Jet::HipTensor.
//
// The cases and literal values are those of the reference's CudaTensor test
// suite (test/Test_CudaTensor.cpp:21-678), run against HipTensor through the
// C ABI.  A small host tensor type stands in for Jet::Tensor in the
// conversion checks (indices, shape and row-major data, the interface the
// conversions use).  hip_tensor_check() returns 0 when every case passes,
// otherwise the number of the first failing case with its message in err.
#include "jet_hip/HipTensor.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using c64 = std::complex<float>;
using c128 = std::complex<double>;
using Jet::HipTensor;

namespace {

template <class U> struct HostTensor {
    using scalar_type_t = U;
    std::vector<std::string> indices;
    std::vector<size_t> shape;
    std::vector<U> data;
    HostTensor(const std::vector<std::string> &i, const std::vector<size_t> &s,
               const std::vector<U> &d)
        : indices(i), shape(s), data(d)
    {
    }
    const std::vector<std::string> &GetIndices() const { return indices; }
    const std::vector<size_t> &GetShape() const { return shape; }
    const std::vector<U> &GetData() const { return data; }
    bool operator==(const HostTensor &o) const
    {
        return indices == o.indices && shape == o.shape && data == o.data;
    }
};

struct Failure {
    std::string what;
};

#define EXPECT(cond)                                                                        \
    do {                                                                                    \
        if (!(cond)) throw Failure{std::string("line ") + std::to_string(__LINE__) + ": " #cond}; \
    } while (0)

template <class F> bool Throws(F &&f, const char *contains)
{
    try {
        f();
    }
    catch (const Jet::Exception &e) {
        return std::string(e.what()).find(contains) != std::string::npos;
    }
    return false;
}

template <class T> bool Close(const std::vector<T> &a, const std::vector<T> &b, double tol)
{
    if (a.size() != b.size()) return false;
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); i++) {
        num += std::norm(std::complex<double>(a[i]) - std::complex<double>(b[i]));
        den += std::norm(std::complex<double>(b[i]));
    }
    return std::sqrt(num) <= tol * std::sqrt(den > 0 ? den : 1.0);
}

template <class T> void CaseConstructors()
{
    HipTensor<T> t0;
    EXPECT(t0.GetSize() == 1 && t0.GetIndices().empty() && t0.GetShape().empty());
    EXPECT(t0.GetHostDataVector() == std::vector<T>{T(0)});

    HipTensor<T> t1(std::vector<size_t>{2, 3});
    EXPECT((t1.GetIndices() == std::vector<std::string>{"?a", "?b"}));
    EXPECT(t1.GetSize() == 6);
    EXPECT(t1.GetHostDataVector() == std::vector<T>(6, T(0)));
    EXPECT(t1.GetIndexToDimension().at("?a") == 2 && t1.GetIndexToDimension().at("?b") == 3);

    const std::vector<T> data{{1, 2}, {3, 4}, {5, 6}, {7, 8}};
    HipTensor<T> t2({"i", "j"}, {2, 2}, data);
    EXPECT(t2.GetHostDataVector() == data);

    HipTensor<T> t3(t2); // copy
    EXPECT(t3 == t2 && t3.GetData() != t2.GetData());
    HipTensor<T> t4;
    t4 = t2; // copy assignment
    EXPECT(t4 == t2);
    HipTensor<T> t5(std::move(t4)); // move
    EXPECT(t5 == t2 && t4.GetData() == nullptr);

    EXPECT(Throws([&] { HipTensor<T>({"i"}, {3}, std::vector<T>{T(1)}); },
                  "Size of data and tensor do not match."));
}

void CaseHostConversion()
{
    // CudaTensor conversion to Tensor (Test_CudaTensor.cpp:376-385).
    HipTensor<c64> dev({"i"}, {2}, std::vector<c64>{{2, 0}, {0, 1}});
    HostTensor<c64> host({"i"}, {2}, {{2, 0}, {0, 1}});
    EXPECT(static_cast<HostTensor<c64>>(dev) == host);

    // Copy from a host tensor: indices and shape reversed, bytes kept
    // (Test_CudaTensor.cpp:328-350).
    HostTensor<c64> h2({"i", "j"}, {2, 3}, {{1, 0}, {2, 0}, {3, 0}, {4, 0}, {5, 0}, {6, 0}});
    HipTensor<c64> d2(h2);
    EXPECT((d2.GetIndices() == std::vector<std::string>{"j", "i"}));
    EXPECT((d2.GetShape() == std::vector<size_t>{3, 2}));
    EXPECT(d2.GetHostDataVector() == h2.data);
    EXPECT(static_cast<HostTensor<c64>>(d2) == h2);
}

void CaseRenameReshape()
{
    // Test_CudaTensor.cpp:387-449.
    HipTensor<c64> t({"a", "b"}, {3, 2});
    t.RenameIndex(1, "z");
    EXPECT((t.GetIndices() == std::vector<std::string>{"a", "z"}));
    EXPECT(t.GetIndexToDimension().at("z") == 2);
    t.RenameIndex(0, "a");
    EXPECT(t.GetIndexToDimension().at("a") == 3);
    EXPECT(Throws([&] { t.RenameIndex(1, "a"); },
                  "Renaming index to already existing value is not allowed."));

    const std::vector<c64> data{{1, 0}, {2, 0}, {3, 0}, {4, 0}, {5, 0}, {6, 0}};
    HipTensor<c64> r({"x", "y"}, {2, 3}, data);
    HipTensor<c64> want({"?a", "?b"}, {3, 2}, data);
    auto got = r.Reshape({3, 2});
    EXPECT(got.GetShape() == want.GetShape() && got.GetIndices() == want.GetIndices());
    EXPECT(got.GetHostDataVector() == data);
    EXPECT(Throws([&] { HipTensor<c64>::Reshape(r, {3, 3}); },
                  "Size is inconsistent between tensors."));
}

void CaseSliceIndex()
{
    // Test_CudaTensor.cpp:451-476 (column-major: x is the fastest index).
    HipTensor<c64> t({"x", "y"}, {2, 3},
                     std::vector<c64>{{1, 0}, {2, 0}, {3, 0}, {4, 0}, {5, 0}, {6, 0}});
    auto host = [](const HipTensor<c64> &h) { return static_cast<HostTensor<c64>>(h); };
    EXPECT(host(t.SliceIndex("x", 0)) == HostTensor<c64>({"y"}, {3}, {{1, 0}, {3, 0}, {5, 0}}));
    EXPECT(host(t.SliceIndex("x", 1)) == HostTensor<c64>({"y"}, {3}, {{2, 0}, {4, 0}, {6, 0}}));
    EXPECT(host(t.SliceIndex("y", 0)) == HostTensor<c64>({"x"}, {2}, {{1, 0}, {2, 0}}));
    EXPECT(host(t.SliceIndex("y", 1)) == HostTensor<c64>({"x"}, {2}, {{3, 0}, {4, 0}}));
    EXPECT(host(t.SliceIndex("y", 2)) == HostTensor<c64>({"x"}, {2}, {{5, 0}, {6, 0}}));
    EXPECT(Throws([&] { t.SliceIndex("q", 0); }, "Sliced index does not exist."));
}

template <class T> void CaseContractLiterals()
{
    // Test_CudaTensor.cpp:478-559.
    {
        HipTensor<T> a({"a", "b"}, {2, 2}, std::vector<T>{{0, 0}, {1, 0}, {1, 0}, {0, 0}});
        HipTensor<T> b({"b"}, {2}, std::vector<T>{{1, 0}, {0, 0}});
        auto c = HipTensor<T>::ContractTensors(a, b);
        EXPECT((c.GetIndices() == std::vector<std::string>{"a"}));
        EXPECT((c.GetHostDataVector() == std::vector<T>{{0, 0}, {1, 0}}));
    }
    {
        HipTensor<T> a({"a", "b", "c"}, {2, 3, 4}, std::vector<T>(24, T(0.5, 0.25)));
        HipTensor<T> b({"b", "c", "d"}, {3, 4, 2}, std::vector<T>(24, T(0.5, 0.25)));
        auto c = a.ContractTensors(b);
        EXPECT((c.GetIndices() == std::vector<std::string>{"a", "d"}));
        EXPECT((c.GetShape() == std::vector<size_t>{2, 2}));
        EXPECT(Close(c.GetHostDataVector(), std::vector<T>(4, T(2.25, 3.0)), 1e-6));
    }
    {
        HipTensor<T> a({"a", "b"}, {2, 2}, std::vector<T>(4, T(1, 0)));
        HipTensor<T> b({"a", "b"}, {2, 2}, std::vector<T>(4, T(1, 0)));
        auto c = HipTensor<T>::ContractTensors(a, b);
        EXPECT(c.GetIndices().empty() && c.GetSize() == 1);
        EXPECT((c.GetHostDataVector() == std::vector<T>{T(4, 0)}));
    }
}

template <class T> void CaseContractRandom(jet_hip_precision_t prec, double tol)
{
    // Test_CudaTensor.cpp:561-626: T1(a,b,c) . T2(c,b,d) against a direct
    // column-major sum.
    HipTensor<T> t1({"a", "b", "c"}, {2, 3, 5});
    HipTensor<T> t2({"c", "b", "d"}, {5, 3, 4});
    t1.FillRandom(7);
    t2.FillRandom(7);
    auto t3 = HipTensor<T>::ContractTensors(t1, t2, prec);
    EXPECT((t3.GetIndices() == std::vector<std::string>{"a", "d"}));
    EXPECT((t3.GetShape() == std::vector<size_t>{2, 4}));
    const auto h1 = t1.GetHostDataVector(), h2 = t2.GetHostDataVector();
    for (auto v : h1) EXPECT(v.real() >= 0 && v.real() < 1 && v.imag() >= 0 && v.imag() < 1);
    std::vector<T> want(8);
    for (int a = 0; a < 2; a++)
        for (int d = 0; d < 4; d++) {
            std::complex<double> s = 0;
            for (int b = 0; b < 3; b++)
                for (int c = 0; c < 5; c++)
                    s += std::complex<double>(h1[a + 2 * b + 6 * c]) *
                         std::complex<double>(h2[c + 5 * b + 15 * d]);
            want[a + 2 * d] = T(s);
        }
    EXPECT(Close(t3.GetHostDataVector(), want, tol));

    // Workspace-explicit path (GetCudaContractionPlan +
    // ContractTensorsWithoutAllocation, CudaTensor.hpp:470-669).
    HipTensor<T> out({"a", "d"}, {2, 4});
    typename HipTensor<T>::HipContractionPlan plan;
    HipTensor<T>::GetHipContractionPlan(plan, t1, t2, out, prec);
    HipTensor<T>::ContractTensorsWithoutAllocation(t1, t2, out, plan);
    EXPECT(Close(out.GetHostDataVector(), want, tol));
}

void CaseAddTensors()
{
    // Test_CudaTensor.cpp:628-677.
    auto host = [](const HipTensor<c64> &h) { return static_cast<HostTensor<c64>>(h); };
    {
        HipTensor<c64> lhs({}, {}, std::vector<c64>{{1, 0}});
        HipTensor<c64> rhs({}, {}, std::vector<c64>{{2, 4}});
        EXPECT(host(HipTensor<c64>::AddTensors(lhs, rhs)) == HostTensor<c64>({}, {}, {{3, 4}}));
    }
    {
        HipTensor<c64> lhs({"i"}, {3}, std::vector<c64>{{1, 0}, {0, 2}, {3, 0}});
        HipTensor<c64> rhs({"i"}, {3}, std::vector<c64>{{1, 0}, {0, 2}, {0, 3}});
        EXPECT(host(lhs.AddTensor(rhs)) ==
               HostTensor<c64>({"i"}, {3}, {{2, 0}, {0, 4}, {3, 3}}));
    }
    {
        HipTensor<c64> lhs({"i", "j"}, {2, 2}, std::vector<c64>{{1, 0}, {2, 0}, {3, 0}, {4, 0}});
        HipTensor<c64> rhs({"i", "j"}, {2, 2}, std::vector<c64>{{0, 1}, {0, 2}, {0, 3}, {0, 4}});
        EXPECT(host(lhs.AddTensor(rhs)) ==
               HostTensor<c64>({"j", "i"}, {2, 2}, {{1, 1}, {2, 2}, {3, 3}, {4, 4}}));
    }
    {
        HipTensor<c64> lhs({"i", "j"}, {2, 2}, std::vector<c64>{{1, 0}, {2, 0}, {3, 0}, {4, 0}});
        HipTensor<c64> rhs({"j", "i"}, {2, 2}, std::vector<c64>{{0, 1}, {0, 2}, {0, 3}, {0, 4}});
        EXPECT(host(HipTensor<c64>::AddTensors(lhs, rhs)) ==
               HostTensor<c64>({"j", "i"}, {2, 2}, {{1, 1}, {2, 3}, {3, 2}, {4, 4}}));
    }
    {
        // The default (zero) tensor is the identity of the reduction.
        HipTensor<c64> zero;
        HipTensor<c64> v({"i"}, {2}, std::vector<c64>{{1, 2}, {3, 4}});
        EXPECT(HipTensor<c64>::AddTensors(zero, v) == v);
        EXPECT(HipTensor<c64>::AddTensors(v, zero) == v);
        HipTensor<c64> w({"k"}, {2});
        EXPECT(Throws([&] { HipTensor<c64>::AddTensors(v, w); },
                      "Tensor addition with disjoint indices is not supported."));
    }
}

void CaseFillTranspose()
{
    // Test_CudaTensor.cpp:180-221: equal seeds give equal data.
    HipTensor<c64> a({"a", "b", "c"}, {3, 2, 3}), b({"a", "b", "c"}, {3, 2, 3});
    a.FillRandom(7);
    b.FillRandom(7);
    EXPECT(a == b);
    b.FillRandom(5);
    EXPECT(a != b);

    // Transpose by labels and by ordering.
    HipTensor<c64> t({"a", "b", "c"}, {2, 3, 4});
    t.FillRandom(11);
    const auto h = t.GetHostDataVector();
    auto u = t.Transpose(std::vector<std::string>{"c", "a", "b"});
    EXPECT((u.GetShape() == std::vector<size_t>{4, 2, 3}));
    const auto hu = u.GetHostDataVector();
    for (int a_ = 0; a_ < 2; a_++)
        for (int b_ = 0; b_ < 3; b_++)
            for (int c_ = 0; c_ < 4; c_++)
                EXPECT(hu[c_ + 4 * a_ + 8 * b_] == h[a_ + 2 * b_ + 6 * c_]);
    auto v = t.Transpose(std::vector<size_t>{2, 0, 1});
    EXPECT(v == u);
}

void CaseErrors()
{
    HipTensor<c64> a({"a", "b"}, {2, 3}), b({"b", "c"}, {4, 2});
    EXPECT(Throws([&] { HipTensor<c64>::ContractTensors(a, b); },
                  "Shared index has different extents."));
}

} // namespace

extern "C" int hip_tensor_check(char *err, int errlen)
{
    struct Case {
        const char *name;
        void (*fn)();
    };
    const Case cases[] = {
        {"constructors c64", CaseConstructors<c64>},
        {"constructors c128", CaseConstructors<c128>},
        {"host conversion", CaseHostConversion},
        {"rename/reshape", CaseRenameReshape},
        {"slice index", CaseSliceIndex},
        {"contract literals c64", CaseContractLiterals<c64>},
        {"contract literals c128", CaseContractLiterals<c128>},
        {"contract random fp32", [] { CaseContractRandom<c64>(JET_HIP_PREC_FP32, 1e-6); }},
        {"contract random bf16x3", [] { CaseContractRandom<c64>(JET_HIP_PREC_BF16X3, 1e-4); }},
        {"contract random fp64", [] { CaseContractRandom<c128>(JET_HIP_PREC_FP64, 1e-14); }},
        {"add tensors", CaseAddTensors},
        {"fill/transpose", CaseFillTranspose},
        {"errors", CaseErrors},
    };
    int i = 0;
    for (const auto &c : cases) {
        i++;
        try {
            c.fn();
        }
        catch (const Failure &f) {
            std::snprintf(err, errlen, "%s: %s", c.name, f.what.c_str());
            return i;
        }
        catch (const std::exception &e) {
            std::snprintf(err, errlen, "%s: exception %s", c.name, e.what());
            return i;
        }
    }
    std::snprintf(err, errlen, "%d cases passed", i);
    return 0;
}
