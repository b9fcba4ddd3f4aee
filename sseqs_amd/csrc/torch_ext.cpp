// PyTorch extension binding of the reference's three extension modules
// (sseqs/xbit_bind.cpp:33-36, sseqs/filter_bind.cpp:67-82,
// sseqs/sw_bind.cpp:111-132) over the C ABI of libsseqs_amd.so.
//
// Same function names, argument names, defaults, dtypes and shapes as the
// reference's pybind11 modules; every launch goes to the current HIP
// stream, and a non-zero return of the C ABI raises (the reference skipped
// unsupported sizes silently).  Host code only: the kernels live in the
// library, which this module links by its run path.
#include <torch/extension.h>

#include <c10/hip/HIPStream.h>

#include "sseqs_amd.h"

namespace {

hipStream_t stream() { return c10::hip::getCurrentHIPStream().stream(); }

void check(int rc, const char* what) {
    TORCH_CHECK(rc == 0, what, " failed (", rc, "): ", sseqs_last_error());
}

torch::Tensor dev_u8(const torch::Tensor& t, const char* name) {
    TORCH_CHECK(t.is_cuda(), name, " must be a device tensor");
    TORCH_CHECK(t.scalar_type() == torch::kUInt8, name, " must be uint8");
    return t.contiguous();
}

torch::Tensor dev_i32(const torch::Tensor& t, const char* name) {
    TORCH_CHECK(t.is_cuda(), name, " must be a device tensor");
    TORCH_CHECK(t.scalar_type() == torch::kInt32, name, " must be int32");
    return t.contiguous();
}

// ---------------------------------------------------------------- xbit --

torch::Tensor compress(torch::Tensor input) {
    auto x = input.contiguous().to(input.is_cuda() ? input.device() : torch::Device(torch::kCUDA),
                                   torch::kUInt8);
    const int64_t n = x.numel();
    auto out = torch::zeros({(n * 5 + 7) / 8}, x.options());
    if (n) check(compress_launcher(x.data_ptr<uint8_t>(), out.data_ptr<uint8_t>(), n, stream()),
                 "compress");
    return out;
}

torch::Tensor decompress(torch::Tensor input, int64_t n, int64_t lo) {
    auto x = input.to(input.is_cuda() ? input.device() : torch::Device(torch::kCUDA),
                      torch::kUInt8).contiguous().view({-1});
    TORCH_CHECK(n >= 0 && lo >= 0, "decompress: n and lo must be >= 0");
    auto out = torch::empty({n}, x.options());
    if (n == 0) return out;
    const int64_t need = lo + ((n + 7) / 8) * 5;  // whole 5-byte groups are read
    if (need > x.numel()) {
        TORCH_CHECK(lo + (n * 5 + 7) / 8 <= x.numel(),
                    "decompress: input holds fewer than n symbols after lo");
        auto padded = torch::zeros({need}, x.options());
        padded.narrow(0, 0, x.numel()).copy_(x);
        x = padded;
    }
    check(decompress_launcher(x.data_ptr<uint8_t>(), out.data_ptr<uint8_t>(), n, lo, stream()),
          "decompress");
    return out;
}

// ---------------------------------------------------------------- conv --

torch::Tensor warp_corr(torch::Tensor corr, torch::Tensor query_reversed_indices,
                        torch::Tensor weight_LUTS_A, torch::Tensor W, torch::Tensor flat,
                        torch::Tensor offsets, torch::Tensor lengths, int64_t k,
                        int64_t query_length, int64_t maxlen, int64_t n,
                        torch::Tensor query_rfft_tensor) {
    (void)query_length;  // taken from query_reversed_indices, as the reference does
    const int64_t B = lengths.size(0);
    const int64_t V = W.size(1);
    auto out = torch::empty({B, 2}, corr.options().dtype(torch::kFloat32));
    // a float32 corr is scored into a temporary fp16 copy (sseqs/filter_bind.cpp:34-35)
    auto corr_half = corr.scalar_type() == torch::kFloat32 ? corr.to(torch::kHalf) : corr;
    TORCH_CHECK(corr_half.scalar_type() == torch::kHalf, "corr must be float16 or float32");
    TORCH_CHECK(corr_half.is_contiguous() && corr_half.numel() >= B,
                "corr must be a contiguous tensor with one entry per segment");
    if (B == 0) return out;
    auto qrev = dev_u8(query_reversed_indices, "query_reversed_indices");
    auto lutA = weight_LUTS_A.contiguous();
    TORCH_CHECK(lutA.scalar_type() == torch::kFloat32 && lutA.is_cuda(),
                "weight_LUTS_A must be a float32 device tensor");
    auto lutB = W.contiguous().view({-1});
    TORCH_CHECK(lutB.scalar_type() == torch::kFloat32 && lutB.is_cuda(),
                "W must be a float32 device tensor");
    if (lutB.numel() <= k * V)  // code 21 reads one entry past rank k - 1
        lutB = torch::cat({lutB, torch::zeros({k * V + 1 - lutB.numel()}, lutB.options())});
    auto f = dev_u8(flat, "flat");
    auto off = dev_i32(offsets, "offsets");
    auto ln = dev_i32(lengths, "lengths");
    check(launch_warp_corr(reinterpret_cast<uint16_t*>(corr_half.data_ptr()),
                           qrev.data_ptr<uint8_t>(), lutA.data_ptr<float>(),
                           lutB.data_ptr<float>(), f.data_ptr<uint8_t>(), off.data_ptr<int32_t>(),
                           ln.data_ptr<int32_t>(), out.data_ptr<float>(), (int)B, (int)V, (int)k,
                           (int)qrev.numel(), (int)maxlen, (int)n, stream(),
                           reinterpret_cast<uint16_t*>(query_rfft_tensor.data_ptr())),
          "warp_corr");
    return out;
}

// ------------------------------------------------------------------ sw --

torch::Tensor sw_cuda_blosum62(torch::Tensor query_indices_tensor, torch::Tensor good_idx,
                               torch::Tensor ascii_tensor, torch::Tensor lengths_tensor,
                               int64_t gap_penalty_param) {
    auto q = dev_u8(query_indices_tensor, "query_indices_tensor");
    auto gi = dev_i32(good_idx, "good_idx");
    auto a = dev_u8(ascii_tensor, "ascii_tensor");
    auto P = dev_i32(lengths_tensor, "lengths_tensor");
    const int64_t M = gi.size(0);
    auto out = torch::empty({M}, gi.options());
    if (M)
        check(launch_sw_cuda_blosum62(q.data_ptr<uint8_t>(), (int)q.size(0), gi.data_ptr<int32_t>(),
                                      a.data_ptr<uint8_t>(), P.data_ptr<int32_t>(), (int)M,
                                      out.data_ptr<int32_t>(), (int)gap_penalty_param, stream()),
              "sw_cuda_blosum62");
    return out;
}

torch::Tensor sw_cuda_affine(torch::Tensor query_indices_tensor, torch::Tensor good_idx,
                             torch::Tensor ascii_tensor, torch::Tensor lengths_tensor,
                             int64_t gap_open, int64_t gap_extend) {
    auto q = dev_u8(query_indices_tensor, "query_indices_tensor");
    auto gi = dev_i32(good_idx, "good_idx");
    auto a = dev_u8(ascii_tensor, "ascii_tensor");
    auto P = dev_i32(lengths_tensor, "lengths_tensor");
    const int64_t M = gi.size(0);
    auto out = torch::empty({M}, gi.options());
    if (M)
        check(launch_sw_cuda_affine(q.data_ptr<uint8_t>(), (int)q.size(0), gi.data_ptr<int32_t>(),
                                    a.data_ptr<uint8_t>(), P.data_ptr<int32_t>(), (int)M,
                                    out.data_ptr<int32_t>(), (int)gap_open, (int)gap_extend,
                                    stream()),
              "sw_cuda_affine");
    return out;
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    auto xbit = m.def_submodule("xbit", "5-bit database codec (sseqs/xbit_bind.cpp)");
    xbit.def("compress", &compress, py::arg("input"));
    xbit.def("decompress", &decompress, py::arg("input"), py::arg("n"), py::arg("lo"));
    auto conv = m.def_submodule("conv", "low-rank correlation prefilter (sseqs/filter_bind.cpp)");
    conv.def("warp_corr", &warp_corr, py::arg("corr"), py::arg("query_reversed_indices"),
             py::arg("weight_LUTS_A"), py::arg("W"), py::arg("flat"), py::arg("offsets"),
             py::arg("lengths"), py::arg("k"), py::arg("query_length"), py::arg("maxlen"),
             py::arg("n"), py::arg("query_rfft_tensor"));
    auto sw = m.def_submodule("sw", "Smith-Waterman scoring (sseqs/sw_bind.cpp)");
    sw.def("sw_cuda_blosum62", &sw_cuda_blosum62, py::arg("query_indices_tensor"),
           py::arg("good_idx"), py::arg("ascii_tensor"), py::arg("lengths_tensor"),
           py::arg("gap_penalty_param") = -10);
    sw.def("sw_cuda_affine", &sw_cuda_affine, py::arg("query_indices_tensor"),
           py::arg("good_idx"), py::arg("ascii_tensor"), py::arg("lengths_tensor"),
           py::arg("gap_open") = 11, py::arg("gap_extend") = 1);
}
