// HipTensor.hpp -- header-only C++ drop-in for Jet::CudaTensor over libjet_hip.
//
// Jet::HipTensor<T, DEVICE> has the interface of Jet::CudaTensor<T, DEVICE>
// (include/jet/CudaTensor.hpp:28-944 of the Jet 0.2.3 tree) and the same
// column-major convention (first index fastest, CudaTensorHelpers.hpp:
// 117-127; conversions from a host Jet::Tensor reverse indices and shape,
// CudaTensor.hpp:312-323, 401-411).  It satisfies the TensorType concept of
// TaskBasedContractor / TensorNetwork / PathInfo (TaskBasedContractor.hpp:
// 23-29, TensorNetwork.hpp:23-37) and calls the GPU only through the C ABI
// of include/jet_hip.h, so it builds with any C++17 compiler and links
// against libjet_hip.so.
//
// T is std::complex<float> or std::complex<double> (the reference's
// cuComplex / cuDoubleComplex have the same layout).  Errors are thrown as
// Jet::Exception carrying the library's message, in the reference's format
// (Abort.hpp:90-97).  Every operation is queued on the device context's
// stream; host reads (GetHostDataVector, conversions) wait for it.
#pragma once

#include "jet_hip.h"

#include <algorithm>
#include <complex>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#if __has_include("jet/Abort.hpp")
// Jet's own exception type, whichever header comes first.
#include "jet/Abort.hpp"
#else
// Without Jet's headers: the same exception type.
namespace Jet {
class Exception : public std::exception {
  public:
    explicit Exception(const std::string &err_msg) noexcept : err_msg(err_msg) {}
    virtual ~Exception() = default;
    const char *what() const noexcept { return err_msg.c_str(); }

  private:
    std::string err_msg;
};
} // namespace Jet
#endif

namespace Jet {

namespace HipTensorHelpers {

inline void Check(int status)
{
    if (status != JET_HIP_SUCCESS) throw Jet::Exception(jet_hip_last_error());
}

[[noreturn]] inline void Fail(const char *msg)
{
    std::ostringstream os;
    os << "[HipTensor.hpp][Method:HipTensor]: Error in Jet: " << msg;
    throw Jet::Exception(os.str());
}

template <class T> constexpr jet_hip_dtype_t DType()
{
    static_assert(std::is_same_v<T, std::complex<float>> ||
                      std::is_same_v<T, std::complex<double>>,
                  "HipTensor supports std::complex<float> and std::complex<double>.");
    return std::is_same_v<T, std::complex<float>> ? JET_HIP_C64 : JET_HIP_C128;
}

// One context (and stream) per device for the whole process.
inline jet_hip_ctx_t Context(int device)
{
    static std::mutex mu;
    static std::map<int, jet_hip_ctx_t> ctxs;
    std::lock_guard<std::mutex> lk(mu);
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second;
    jet_hip_ctx_t c = nullptr;
    Check(jet_hip_ctx_create(device, &c));
    ctxs[device] = c;
    return c;
}

inline size_t ShapeToSize(const std::vector<size_t> &shape)
{
    size_t n = 1;
    for (auto s : shape) n *= s;
    return n;
}

template <class V> V Reversed(V v)
{
    std::reverse(v.begin(), v.end());
    return v;
}

// Integer modes for label lists (shared ids for equal labels).
struct Modes {
    std::unordered_map<std::string, int32_t> ids;
    std::vector<int32_t> operator()(const std::vector<std::string> &labels)
    {
        std::vector<int32_t> out;
        for (const auto &l : labels) {
            auto it = ids.find(l);
            if (it == ids.end()) it = ids.emplace(l, (int32_t)ids.size()).first;
            out.push_back(it->second);
        }
        return out;
    }
};

inline std::vector<std::string> Subtract(const std::vector<std::string> &a,
                                         const std::vector<std::string> &b)
{
    std::vector<std::string> out;
    for (const auto &x : a)
        if (std::find(b.begin(), b.end(), x) == b.end()) out.push_back(x);
    return out;
}

} // namespace HipTensorHelpers

template <class T = std::complex<float>, int HIP_DEVICE = 0> class HipTensor {
  public:
    using scalar_type_t = T;
    using scalar_type_t_precision = typename T::value_type;

    // --- construction (CudaTensor.hpp:145-214) ---------------------------
    HipTensor() : data_(nullptr)
    {
        InitIndicesAndShape({}, {});
        const T zero{};
        Upload(&zero);
    }

    HipTensor(const std::vector<std::string> &indices, const std::vector<size_t> &shape)
        : data_(nullptr)
    {
        InitIndicesAndShape(indices, shape);
        HipTensorHelpers::Check(
            jet_hip_memset(Ctx(), data_, 0, sizeof(T) * GetSize(), nullptr));
    }

    HipTensor(const std::vector<std::string> &indices, const std::vector<size_t> &shape,
              const std::vector<T> &data)
        : HipTensor(indices, shape)
    {
        if (data.size() != GetSize())
            HipTensorHelpers::Fail("Size of data and tensor do not match.");
        Upload(data.data());
    }

    HipTensor(const std::vector<std::string> &indices, const std::vector<size_t> &shape,
              const T *data)
        : HipTensor(indices, shape)
    {
        Upload(data);
    }

    explicit HipTensor(const std::vector<size_t> &shape) : data_(nullptr)
    {
        static const char *alpha = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ";
        std::vector<std::string> idx(shape.size());
        for (size_t i = 0; i < shape.size(); i++) {
            const size_t div = i / 52;
            idx[i] = std::string("?") + alpha[i % 52] + (div ? std::to_string(div - 1) : "");
        }
        InitIndicesAndShape(idx, shape);
        HipTensorHelpers::Check(
            jet_hip_memset(Ctx(), data_, 0, sizeof(T) * GetSize(), nullptr));
    }

    HipTensor(const HipTensor &other) : data_(nullptr)
    {
        InitIndicesAndShape(other.indices_, other.shape_);
        HipTensorHelpers::Check(jet_hip_memcpy(Ctx(), data_, other.data_,
                                               sizeof(T) * GetSize(), 2, nullptr, 1));
    }

    HipTensor(HipTensor &&other) noexcept
        : data_(other.data_), indices_(std::move(other.indices_)),
          shape_(std::move(other.shape_)),
          index_to_dimension_(std::move(other.index_to_dimension_))
    {
        other.data_ = nullptr;
    }

    // From a host tensor with GetIndices / GetShape / GetData (Jet::Tensor):
    // indices and shape reversed, same bytes (CudaTensor.hpp:312-323).
    template <class HostTensor,
              class = decltype(std::declval<const HostTensor &>().GetData().data())>
    explicit HipTensor(const HostTensor &host) : data_(nullptr)
    {
        static_assert(sizeof(typename HostTensor::scalar_type_t) == sizeof(T),
                      "Size of CPU and GPU data types do not match.");
        InitIndicesAndShape(HipTensorHelpers::Reversed(host.GetIndices()),
                            HipTensorHelpers::Reversed(host.GetShape()));
        Upload(reinterpret_cast<const T *>(host.GetData().data()));
    }

    ~HipTensor() { Release(); }

    // To a host tensor type constructible from (indices, shape, data), such
    // as Jet::Tensor: indices and shape reversed, same bytes
    // (CudaTensor.hpp:401-411).
    template <class HostTensor,
              class = decltype(HostTensor(std::vector<std::string>{}, std::vector<size_t>{},
                                          std::vector<typename HostTensor::scalar_type_t>{}))>
    explicit operator HostTensor() const
    {
        using H = typename HostTensor::scalar_type_t;
        static_assert(sizeof(H) == sizeof(T), "Size of CPU and GPU data types do not match.");
        std::vector<H> host(GetSize());
        CopyGpuDataToHost(reinterpret_cast<T *>(host.data()));
        return HostTensor(HipTensorHelpers::Reversed(indices_),
                          HipTensorHelpers::Reversed(shape_), host);
    }

    HipTensor &operator=(const HipTensor &other)
    {
        if (this != &other) {
            InitIndicesAndShape(other.indices_, other.shape_);
            HipTensorHelpers::Check(jet_hip_memcpy(Ctx(), data_, other.data_,
                                                   sizeof(T) * GetSize(), 2, nullptr, 1));
        }
        return *this;
    }

    HipTensor &operator=(HipTensor &&other) noexcept
    {
        if (this != &other) {
            Release();
            data_ = other.data_;
            other.data_ = nullptr;
            indices_ = std::move(other.indices_);
            shape_ = std::move(other.shape_);
            index_to_dimension_ = std::move(other.index_to_dimension_);
        }
        return *this;
    }

    void InitIndicesAndShape(const std::vector<std::string> &indices,
                             const std::vector<size_t> &shape)
    {
        Release();
        indices_ = indices;
        shape_ = shape;
        index_to_dimension_.clear();
        for (size_t i = 0; i < shape_.size(); i++) index_to_dimension_[indices_[i]] = shape_[i];
        void *p = nullptr;
        HipTensorHelpers::Check(jet_hip_malloc(Ctx(), sizeof(T) * GetSize(), &p));
        data_ = static_cast<T *>(p);
    }

    // --- accessors --------------------------------------------------------
    const std::vector<std::string> &GetIndices() const { return indices_; }
    const std::vector<size_t> &GetShape() const { return shape_; }
    size_t GetSize() const { return HipTensorHelpers::ShapeToSize(shape_); }
    T *GetData() { return data_; }
    const T *GetData() const { return data_; }
    const std::unordered_map<std::string, size_t> &GetIndexToDimension() const
    {
        return index_to_dimension_;
    }

    void RenameIndex(size_t ind, std::string new_string)
    {
        const std::string old_string = indices_[ind];
        if (old_string == new_string) return;
        if (index_to_dimension_.count(new_string))
            HipTensorHelpers::Fail("Renaming index to already existing value is not allowed.");
        indices_[ind] = new_string;
        index_to_dimension_[new_string] = index_to_dimension_[old_string];
        index_to_dimension_.erase(old_string);
    }

    std::vector<T> GetHostDataVector() const
    {
        std::vector<T> host(GetSize());
        HipTensorHelpers::Check(
            jet_hip_memcpy(Ctx(), host.data(), data_, sizeof(T) * GetSize(), 1, nullptr, 0));
        return host;
    }

    void CopyHostDataToGpu(const T *host)
    {
        Upload(host);
    }
    void CopyGpuDataToHost(T *host) const
    {
        HipTensorHelpers::Check(
            jet_hip_memcpy(Ctx(), host, data_, sizeof(T) * GetSize(), 1, nullptr, 0));
    }
    void AsyncCopyHostDataToGpu(const T *host, void *stream = nullptr)
    {
        HipTensorHelpers::Check(
            jet_hip_memcpy(Ctx(), data_, host, sizeof(T) * GetSize(), 0, stream, 1));
    }
    void AsyncCopyGpuDataToHost(T *host, void *stream = nullptr) const
    {
        HipTensorHelpers::Check(
            jet_hip_memcpy(Ctx(), host, data_, sizeof(T) * GetSize(), 1, stream, 1));
    }

    // FillRandom(seed): counter-based uniform [0, 1) per component (the
    // range of curandGenerateUniform in CudaTensor.hpp:419-446; the values
    // themselves differ from cuRAND's).
    void FillRandom(size_t seed)
    {
        HipTensorHelpers::Check(jet_hip_fill_random(Ctx(), HipTensorHelpers::DType<T>(),
                                                    GetSize(), seed, 0, data_, nullptr));
    }

    bool operator==(const HipTensor &other) const
    {
        return shape_ == other.shape_ && indices_ == other.indices_ &&
               index_to_dimension_ == other.index_to_dimension_ &&
               GetHostDataVector() == other.GetHostDataVector();
    }
    bool operator!=(const HipTensor &other) const { return !(*this == other); }

    // --- contraction plans (CudaTensor.hpp:470-669) -----------------------
    struct HipContractionPlan {
        jet_hip_plan_t plan = nullptr;
        void *work = nullptr;
        size_t work_size = 0;
        HipContractionPlan() = default;
        HipContractionPlan(const HipContractionPlan &) = delete;
        HipContractionPlan &operator=(const HipContractionPlan &) = delete;
        ~HipContractionPlan()
        {
            if (work) jet_hip_free(HipTensor::Ctx(), work);
            if (plan) jet_hip_plan_destroy(plan);
        }
    };

    static void GetHipContractionPlan(HipContractionPlan &cplan, const HipTensor &a,
                                      const HipTensor &b, const HipTensor &c,
                                      jet_hip_precision_t precision = JET_HIP_PREC_DEFAULT)
    {
        HipTensorHelpers::Modes modes;
        const auto ma = modes(a.indices_), mb = modes(b.indices_), mc = modes(c.indices_);
        std::vector<int64_t> ea(a.shape_.begin(), a.shape_.end());
        std::vector<int64_t> eb(b.shape_.begin(), b.shape_.end());
        HipTensorHelpers::Check(jet_hip_contraction_plan_create(
            Ctx(), HipTensorHelpers::DType<T>(), precision, JET_HIP_COL_MAJOR, (int)ma.size(),
            ma.data(), ea.data(), (int)mb.size(), mb.data(), eb.data(), (int)mc.size(),
            mc.data(), &cplan.plan));
        HipTensorHelpers::Check(jet_hip_plan_workspace_size(cplan.plan, &cplan.work_size));
        if (cplan.work_size)
            HipTensorHelpers::Check(jet_hip_malloc(Ctx(), cplan.work_size, &cplan.work));
    }

    static void ContractTensorsWithoutAllocation(const HipTensor &a, const HipTensor &b,
                                                 HipTensor &c, HipContractionPlan &cplan,
                                                 void *stream = nullptr)
    {
        HipTensorHelpers::Check(jet_hip_contract(cplan.plan, a.data_, b.data_, c.data_,
                                                 cplan.work, cplan.work_size, stream));
    }

    // CudaTensor::ContractTensors (CudaTensor.hpp:216-255): the result's
    // indices are (a \ b) ++ (b \ a), column-major.
    static HipTensor ContractTensors(const HipTensor &a, const HipTensor &b)
    {
        return ContractTensors(a, b, JET_HIP_PREC_DEFAULT);
    }

    static HipTensor ContractTensors(const HipTensor &a, const HipTensor &b,
                                     jet_hip_precision_t precision)
    {
        const auto left = HipTensorHelpers::Subtract(a.indices_, b.indices_);
        const auto right = HipTensorHelpers::Subtract(b.indices_, a.indices_);
        std::vector<std::string> ci = left;
        ci.insert(ci.end(), right.begin(), right.end());
        std::vector<size_t> cs;
        for (const auto &l : left) cs.push_back(a.index_to_dimension_.at(l));
        for (const auto &r : right) cs.push_back(b.index_to_dimension_.at(r));
        HipTensor c(ci, cs);
        auto &plan = CachedPlan(a, b, c, precision);
        HipTensorHelpers::Check(
            jet_hip_contract(plan.plan, a.data_, b.data_, c.data_, nullptr, 0, nullptr));
        return c;
    }

    HipTensor ContractTensors(const HipTensor &other) const
    {
        return ContractTensors(*this, other);
    }

    // --- transpose / add / slice / reshape (CudaTensor.hpp:38-137, 671-905)
    static HipTensor Transpose(const HipTensor &t, const std::vector<std::string> &new_indices)
    {
        if (t.indices_ == new_indices) return t;
        if (t.indices_.empty()) HipTensorHelpers::Fail("Number of indices cannot be zero.");
        std::vector<size_t> shape;
        for (const auto &l : new_indices) shape.push_back(t.index_to_dimension_.at(l));
        HipTensor out(new_indices, shape);
        HipTensorHelpers::Modes modes;
        const auto mi = modes(t.indices_), mo = modes(new_indices);
        std::vector<int64_t> ext(t.shape_.begin(), t.shape_.end());
        HipTensorHelpers::Check(jet_hip_permute(Ctx(), HipTensorHelpers::DType<T>(),
                                                JET_HIP_COL_MAJOR, (int)mi.size(), mi.data(),
                                                ext.data(), mo.data(), t.data_, out.data_,
                                                nullptr));
        return out;
    }

    static HipTensor Transpose(const HipTensor &t, const std::vector<size_t> &new_ordering)
    {
        if (t.indices_.size() != new_ordering.size())
            HipTensorHelpers::Fail("Size of ordering must match number of tensor indices.");
        std::vector<std::string> idx;
        for (auto o : new_ordering) idx.push_back(t.indices_[o]);
        return Transpose(t, idx);
    }

    HipTensor Transpose(const std::vector<std::string> &new_indices) const
    {
        return Transpose(*this, new_indices);
    }
    HipTensor Transpose(const std::vector<size_t> &new_ordering) const
    {
        return Transpose(*this, new_ordering);
    }

    static HipTensor AddTensors(const HipTensor &A, const HipTensor &B)
    {
        static const HipTensor zero;
        if (A == zero) return B;
        if (B == zero) return A;
        if (!HipTensorHelpers::Subtract(A.indices_, B.indices_).empty() ||
            !HipTensorHelpers::Subtract(B.indices_, A.indices_).empty())
            HipTensorHelpers::Fail("Tensor addition with disjoint indices is not supported.");
        HipTensor C(A.indices_, A.shape_);
        HipTensorHelpers::Modes modes;
        const auto ma = modes(A.indices_), mb = modes(B.indices_);
        std::vector<int64_t> ext(A.shape_.begin(), A.shape_.end());
        HipTensorHelpers::Check(jet_hip_add(Ctx(), HipTensorHelpers::DType<T>(),
                                            JET_HIP_COL_MAJOR, (int)ma.size(), ma.data(),
                                            ext.data(), mb.data(), A.data_, B.data_, C.data_,
                                            nullptr));
        return C;
    }

    HipTensor AddTensor(const HipTensor &other) const { return AddTensors(*this, other); }

    static HipTensor SliceIndex(const HipTensor &t, const std::string &index, size_t value)
    {
        auto it = std::find(t.indices_.begin(), t.indices_.end(), index);
        if (it == t.indices_.end()) HipTensorHelpers::Fail("Sliced index does not exist.");
        const int axis = (int)(it - t.indices_.begin());
        std::vector<std::string> idx;
        std::vector<size_t> shape;
        for (size_t i = 0; i < t.indices_.size(); i++)
            if ((int)i != axis) {
                idx.push_back(t.indices_[i]);
                shape.push_back(t.shape_[i]);
            }
        HipTensor out(idx, shape);
        std::vector<int64_t> ext(t.shape_.begin(), t.shape_.end());
        HipTensorHelpers::Check(jet_hip_slice(Ctx(), HipTensorHelpers::DType<T>(),
                                              JET_HIP_COL_MAJOR, (int)ext.size(), ext.data(),
                                              axis, (int64_t)value, t.data_, out.data_,
                                              nullptr));
        return out;
    }

    HipTensor SliceIndex(const std::string &index, size_t value) const
    {
        return SliceIndex(*this, index, value);
    }

    static HipTensor Reshape(const HipTensor &old, const std::vector<size_t> &new_shape)
    {
        if (old.GetSize() != HipTensorHelpers::ShapeToSize(new_shape))
            HipTensorHelpers::Fail("Size is inconsistent between tensors.");
        HipTensor out(new_shape);
        HipTensorHelpers::Check(jet_hip_memcpy(Ctx(), out.data_, old.data_,
                                               sizeof(T) * old.GetSize(), 2, nullptr, 1));
        return out;
    }

    HipTensor Reshape(const std::vector<size_t> &new_shape) const
    {
        return Reshape(*this, new_shape);
    }

    static jet_hip_ctx_t Ctx() { return HipTensorHelpers::Context(HIP_DEVICE); }

  private:
    T *data_;
    std::vector<std::string> indices_;
    std::vector<size_t> shape_;
    std::unordered_map<std::string, size_t> index_to_dimension_;

    void Upload(const T *host)
    {
        HipTensorHelpers::Check(
            jet_hip_memcpy(Ctx(), data_, host, sizeof(T) * GetSize(), 0, nullptr, 0));
    }

    void Release() noexcept
    {
        if (data_) jet_hip_free(Ctx(), data_);
        data_ = nullptr;
    }

    // Plans are reused for contractions of the same (relabelled) problem.
    static HipContractionPlan &CachedPlan(const HipTensor &a, const HipTensor &b,
                                          const HipTensor &c, jet_hip_precision_t precision)
    {
        HipTensorHelpers::Modes modes;
        const auto ma = modes(a.indices_), mb = modes(b.indices_), mc = modes(c.indices_);
        using Key = std::tuple<std::vector<int32_t>, std::vector<size_t>, std::vector<int32_t>,
                               std::vector<size_t>, std::vector<int32_t>, int>;
        static std::mutex mu;
        static std::map<Key, std::unique_ptr<HipContractionPlan>> cache;
        Key key{ma, a.shape_, mb, b.shape_, mc, (int)precision};
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return *it->second;
        auto p = std::make_unique<HipContractionPlan>();
        HipTensorHelpers::Modes m2;
        const auto a2 = m2(a.indices_), b2 = m2(b.indices_), c2 = m2(c.indices_);
        std::vector<int64_t> ea(a.shape_.begin(), a.shape_.end());
        std::vector<int64_t> eb(b.shape_.begin(), b.shape_.end());
        HipTensorHelpers::Check(jet_hip_contraction_plan_create(
            Ctx(), HipTensorHelpers::DType<T>(), precision, JET_HIP_COL_MAJOR, (int)a2.size(),
            a2.data(), ea.data(), (int)b2.size(), b2.data(), eb.data(), (int)c2.size(),
            c2.data(), &p->plan));
        auto &ref = *p;
        cache.emplace(std::move(key), std::move(p));
        return ref;
    }
};

} // namespace Jet
