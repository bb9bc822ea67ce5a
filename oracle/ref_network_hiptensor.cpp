// ref_network_hiptensor.cpp -- the reference's own TensorNetwork and PathInfo
// templates instantiated with Jet::HipTensor (TEST INFRASTRUCTURE ONLY).
//
// Built by `make -C oracle ref` against the reference's headers where they
// lie (TensorNetwork.hpp, PathInfo.hpp, Abort.hpp, Utilities.hpp -- all
// header-only) and libjet_hip; nothing of the reference is copied.  It shows
// that HipTensor satisfies the TensorType concept of the reference's classes
// (TensorNetwork.hpp:23-37, PathInfo.hpp:80-112) and reproduces the values
// the reference's own tests expect of TensorNetwork<CudaTensor>
// (test/Test_TensorNetwork.cpp:649-1215).  jr_network_hiptensor_check()
// returns 0 when every case passes, else the number of the failing case
// with a message in err.  Only tests/ load this library.
#include "jet/PathInfo.hpp"
#include "jet/TensorNetwork.hpp"
#include "jet_hip/HipTensor.hpp"

#include <complex>
#include <cstdio>
#include <string>
#include <vector>

namespace {

using c64 = std::complex<float>;
using T = Jet::HipTensor<c64>;
using Indices = std::vector<std::string>;
using Shape = std::vector<size_t>;

struct Failure {
    std::string what;
};

#define EXPECT(cond)                                                                         \
    do {                                                                                     \
        if (!(cond)) throw Failure{std::string("line ") + std::to_string(__LINE__) + ": " #cond}; \
    } while (0)

// MakeCudaTensor of the reference tests: element i is (i, 2i) in memory order.
T Make(const Indices &idx, const Shape &shape)
{
    T t(idx, shape);
    std::vector<c64> h(t.GetSize());
    for (size_t i = 0; i < h.size(); i++) h[i] = c64(float(i), float(2 * i));
    t.CopyHostDataToGpu(h.data());
    return t;
}

template <class V> V Rev(V v)
{
    std::reverse(v.begin(), v.end());
    return v;
}

void CaseSlice()
{
    // Test_TensorNetwork.cpp:762-915, "Slice [1, :, 2]" and "Slice [1, 2, 3]"
    // (memory order is column-major: A0 fastest).
    for (int which = 0; which < 2; which++) {
        Jet::TensorNetwork<T> tn;
        tn.AddTensor(Make({"A0", "B1", "C2"}, {2, 3, 4}), {});
        tn.AddTensor(Make({"D3"}, {2}), {});
        if (which == 0) {
            tn.SliceIndices({"A0", "C2"}, 1 * 4 + 2);
            const auto &node = tn.GetNodes().front();
            EXPECT(node.name == "A0(1)B1C2(2)");
            EXPECT((node.tensor.GetIndices() == Indices{"B1"}));
            EXPECT((node.tensor.GetShape() == Shape{3}));
            // column-major: element (a=1, b, c=2) sits at 1 + 2b + 6*2
            const auto d = node.tensor.GetHostDataVector();
            for (int b = 0; b < 3; b++) EXPECT(d[b] == c64(13 + 2 * b, 2 * (13 + 2 * b)));
        }
        else {
            tn.SliceIndices({"A0", "B1", "C2"}, 1 * 3 * 4 + 2 * 4 + 3);
            const auto &node = tn.GetNodes().front();
            EXPECT(node.name == "A0(1)B1(2)C2(3)");
            EXPECT(node.tensor.GetIndices().empty());
            EXPECT(node.tensor.GetHostDataVector() == std::vector<c64>{c64(23, 46)});
        }
    }
}

void CaseContract()
{
    // Test_TensorNetwork.cpp:1112-1178: explicit contraction.
    Jet::TensorNetwork<T> tn;
    tn.AddTensor(Make(Rev(Indices{"A0", "B1"}), Rev(Shape{2, 3})), {});
    tn.AddTensor(Make(Rev(Indices{"C2", "B1"}), Rev(Shape{2, 3})), {});
    tn.AddTensor(Make(Rev(Indices{"C2", "D3"}), Rev(Shape{2, 2})), {});
    Jet::PathInfo info(tn, {{1, 2}, {0, 3}});
    const auto result = tn.Contract({{1, 2}, {0, 3}});
    EXPECT((tn.GetPath() == Jet::TensorNetwork<T>::Path{{1, 2}, {0, 3}}));
    EXPECT((result.GetShape() == Shape{2, 2}));
    // row-major A0, D3 order of the column-major result
    const auto t = result.Transpose(Indices{"D3", "A0"});
    const auto d = t.GetHostDataVector();
    const std::vector<c64> want{{-308, -56}, {-517, -94}, {-1100, -200}, {-1804, -328}};
    EXPECT(d == want);
    const auto &nodes = tn.GetNodes();
    EXPECT(nodes.size() == 5 && nodes[3].name == "B1D3" && nodes[4].name == "A0D3");
    EXPECT(nodes[3].contracted && !nodes[4].contracted);
    // PathInfo on the same network: flops |out| * 2 * prod(contracted)
    EXPECT(info.GetNumLeaves() == 3);
    // step 3: B1 (3) x D3 (2) out, C2 (2) contracted; step 4: A0 x D3 out, B1 contracted
    EXPECT(info.GetPathStepFlops(3) == (3.0 * 2.0) * 2.0 * 2.0);
    EXPECT(info.GetPathStepFlops(4) == (2.0 * 2.0) * 2.0 * 3.0);
}

void CaseImplicit()
{
    // Test_TensorNetwork.cpp:1022-1110: implicit contraction [2, 3] -(1)- [3, 3].
    Jet::TensorNetwork<T> tn;
    tn.AddTensor(Make({"A0", "B1"}, {2, 3}), {});
    tn.AddTensor(Make({"C2", "B1"}, {3, 3}), {});
    const auto result = tn.Contract();
    EXPECT((tn.GetPath() == Jet::TensorNetwork<T>::Path{{0, 1}}));
    EXPECT((result.GetShape() == Shape{2, 3}));
    // C(a, c) = sum_b A(a, b) B(c, b) with column-major A(a,b) = a + 2b,
    // B(c,b) = c + 3b, each value v stored as v(1 + 2i).
    const auto d = result.GetHostDataVector();
    for (int c = 0; c < 3; c++)
        for (int a = 0; a < 2; a++) {
            std::complex<double> s = 0;
            for (int b = 0; b < 3; b++)
                s += std::complex<double>(a + 2 * b, 2 * (a + 2 * b)) *
                     std::complex<double>(c + 3 * b, 2 * (c + 3 * b));
            EXPECT(std::complex<double>(d[a + 2 * c]) == s);
        }
}

void CaseErrors()
{
    Jet::TensorNetwork<T> tn;
    bool threw = false;
    try {
        tn.Contract();
    }
    catch (const Jet::Exception &e) {
        threw = std::string(e.what()).find("An empty tensor network cannot be contracted.") !=
                std::string::npos;
    }
    EXPECT(threw);
}

} // namespace

extern "C" int jr_network_hiptensor_check(char *err, int errlen)
{
    struct Case {
        const char *name;
        void (*fn)();
    };
    const Case cases[] = {{"slice", CaseSlice},
                          {"explicit contract + PathInfo", CaseContract},
                          {"implicit contract", CaseImplicit},
                          {"errors", CaseErrors}};
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
