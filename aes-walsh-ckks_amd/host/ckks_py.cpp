// Python binding of the C++ host API (tests and the benchmark drive the
// library through it).  Device-resident objects stay on the device; polys,
// plaintexts and keys cross into numpy only through explicit downloads.
#include <cstring>

#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <chrono>
#include <cstring>

#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "ckks.hpp"

namespace py = pybind11;

namespace {

py::array_t<uint64_t> to_np(const std::vector<uint64_t>& v) {
    py::array_t<uint64_t> a(v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

std::vector<uint64_t> from_np(py::array_t<uint64_t, py::array::c_style | py::array::forcecast> a) {
    return std::vector<uint64_t>(a.data(), a.data() + a.size());
}

MyVector<Complex> complex_vec(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> a) {
    return MyVector<Complex>(a.data(), a.data() + a.size());
}

py::array_t<std::complex<double>> complex_np(const MyVector<Complex>& v) {
    py::array_t<std::complex<double>> a(v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

}  // namespace

// The 1408 encrypted round-key bits, held on the C++ side so that each AES
// call reads them without copying them into Python objects.
struct AesRoundKeyBits {
    explicit AesRoundKeyBits(MyVector<CKKSCiphertext> b) : bits(std::move(b)) {}
    AesRoundKeyBits(const AesRoundKeyBits&) = delete;
    AesRoundKeyBits(AesRoundKeyBits&&) = default;
    MyVector<CKKSCiphertext> bits;
};

std::vector<std::array<uint8_t, 16>> blocks_from_array(const py::array_t<uint8_t, py::array::c_style>& a) {
    if (a.ndim() != 2 || a.shape(1) != 16) throw std::invalid_argument("blocks must be a uint8 array of shape [n, 16]");
    std::vector<std::array<uint8_t, 16>> out((size_t)a.shape(0));
    for (size_t i = 0; i < out.size(); ++i) std::memcpy(out[i].data(), a.data(i, 0), 16);
    return out;
}

PYBIND11_MODULE(ckks_py, m) {
    m.doc() = "C++ host API of the MI355X CKKS library";

    py::enum_<CKKSSecretOwner>(m, "CKKSSecretOwner")
        .value("Dense", CKKSSecretOwner::Dense)
        .value("BootstrapSparse", CKKSSecretOwner::BootstrapSparse);
    py::enum_<CKKSMessageEncodingState>(m, "CKKSMessageEncodingState")
        .value("Slots", CKKSMessageEncodingState::Slots)
        .value("Coefficients", CKKSMessageEncodingState::Coefficients);
    py::enum_<CKKSDecryptDomain>(m, "CKKSDecryptDomain")
        .value("Slots", CKKSDecryptDomain::Slots)
        .value("Coefficients", CKKSDecryptDomain::Coefficients);

    py::class_<CKKSBootstrapParams>(m, "CKKSBootstrapParams")
        .def(py::init<>())
        .def_readwrite("sparse_secret_hamming_weight", &CKKSBootstrapParams::sparse_secret_hamming_weight)
        .def_readwrite("active_slot_count", &CKKSBootstrapParams::active_slot_count)
        .def_readwrite("linear_transform_depths", &CKKSBootstrapParams::linear_transform_depths)
        .def_readwrite("linear_transform_log_plaintext_scales", &CKKSBootstrapParams::linear_transform_log_plaintext_scales)
        .def_readwrite("evalmod_depth", &CKKSBootstrapParams::evalmod_depth)
        .def_readwrite("evalmod_bit_size", &CKKSBootstrapParams::evalmod_bit_size)
        .def_readwrite("evalmod_k", &CKKSBootstrapParams::evalmod_k)
        .def_property("evalmod_k_failure_probability_log2",
                      [](const CKKSBootstrapParams& b) { return (double)b.evalmod_k_failure_probability_log2; },
                      [](CKKSBootstrapParams& b, double v) { b.evalmod_k_failure_probability_log2 = v; })
        .def_readwrite("evalmod_double_angle", &CKKSBootstrapParams::evalmod_double_angle)
        .def_readwrite("linear_transform_bsgs_baby_step_counts", &CKKSBootstrapParams::linear_transform_bsgs_baby_step_counts)
        .def_readwrite("linear_transform_log_bsgs_ratios", &CKKSBootstrapParams::linear_transform_log_bsgs_ratios)
        .def_readwrite("linear_transform_merge_depths", &CKKSBootstrapParams::linear_transform_merge_depths)
        .def_readwrite("linear_transform_use_qp", &CKKSBootstrapParams::linear_transform_use_qp)
        .def_readwrite("evalmod_polynomial_degree", &CKKSBootstrapParams::evalmod_polynomial_degree)
        .def_readwrite("evalmod_log_message_ratio", &CKKSBootstrapParams::evalmod_log_message_ratio);

    py::class_<CKKSAesWalshParams>(m, "CKKSAesWalshParams")
        .def(py::init<>())
        .def_readwrite("enabled", &CKKSAesWalshParams::enabled)
        .def_readwrite("block_count", &CKKSAesWalshParams::block_count)
        .def_readwrite("packed_segment_count", &CKKSAesWalshParams::packed_segment_count);

    py::class_<CKKSParamsConfig>(m, "CKKSParamsConfig")
        .def_readwrite("bts_keys", &CKKSParamsConfig::bts_keys)
        .def_readwrite("bootstrap_params", &CKKSParamsConfig::bootstrap_params)
        .def(py::init<>())
        .def_readwrite("scale", &CKKSParamsConfig::scale)
        .def_readwrite("sigma", &CKKSParamsConfig::sigma)
        .def_readwrite("N", &CKKSParamsConfig::N)
        .def_readwrite("mult_depth", &CKKSParamsConfig::mult_depth)
        .def_readwrite("q0_bit", &CKKSParamsConfig::q0_bit)
        .def_readwrite("q_bit", &CKKSParamsConfig::q_bit)
        .def_readwrite("q_mult_bits", &CKKSParamsConfig::q_mult_bits)
        .def_readwrite("dnum", &CKKSParamsConfig::dnum)
        .def_readwrite("p_bit", &CKKSParamsConfig::p_bit)
        .def_readwrite("pmoduli_count", &CKKSParamsConfig::pmoduli_count)
        .def_readwrite("p_bits", &CKKSParamsConfig::p_bits)
        .def_readwrite("secret_key_hamming_weight", &CKKSParamsConfig::secret_key_hamming_weight)
        .def_readwrite("q_bits", &CKKSParamsConfig::q_bits)
        .def_readwrite("rotate_values", &CKKSParamsConfig::rotate_values)
        .def_readwrite("secure", &CKKSParamsConfig::secure)
        .def_readwrite("aes_walsh_params", &CKKSParamsConfig::aes_walsh_params)
        .def_readwrite("fold_hybrid_keyswitch", &CKKSParamsConfig::fold_hybrid_keyswitch)
        .def_readwrite("montgomery", &CKKSParamsConfig::montgomery)
        .def_readwrite("seed", &CKKSParamsConfig::seed);

    m.def("aes_walsh15_config", &aes_walsh15_config);
    m.def("aes_walsh15_bootstrap_params", &aes_walsh15_bootstrap_params);
    py::class_<CKKSParams>(m, "CKKSParams")
        .def(py::init<CKKSParamsConfig>())
        .def_static("AESWalsh15", &CKKSParams::AESWalsh15)
        .def_static("XBootParamAes13Reference", &CKKSParams::XBootParamAes13Reference)
        .def_static("AESXBoot14Secure", &CKKSParams::AESXBoot14Secure)
        .def_static("AESXBoot14SecureSparse1024", &CKKSParams::AESXBoot14SecureSparse1024)
        .def_static("AESXBoot15", &CKKSParams::AESXBoot15)
        .def_static("AESXBoot15Sparse1024", &CKKSParams::AESXBoot15Sparse1024)
        .def("usesAesWalsh", &CKKSParams::usesAesWalsh)
        .def("getAesWalshBlockCount", &CKKSParams::getAesWalshBlockCount)
        .def("getAesWalshPackedSegmentCount", &CKKSParams::getAesWalshPackedSegmentCount)
        .def("getN", &CKKSParams::getN)
        .def("getSeed", &CKKSParams::getSeed)
        .def("getSlots", &CKKSParams::getSlots)
        .def("getMaxLevel", &CKKSParams::getMaxLevel)
        .def("getDnum", &CKKSParams::getDnum)
        .def("getAlpha", &CKKSParams::getAlpha)
        .def("getScale", &CKKSParams::getScale)
        .def("getSigma", &CKKSParams::getSigma)
        .def("getModuli", [](const CKKSParams& p) { return to_np(p.getModuli()); })
        .def("getPModuli", [](const CKKSParams& p) { return to_np(p.getPModuli()); })
        .def("getDepth", &CKKSParams::getDepth)
        .def("getBootstrapS2CDepth", &CKKSParams::getBootstrapS2CDepth)
        .def("getBootstrapC2SDepth", &CKKSParams::getBootstrapC2SDepth)
        .def("getBootstrapActiveSlotCount", &CKKSParams::getBootstrapActiveSlotCount)
        .def("getBootstrapEvalModK", &CKKSParams::getBootstrapEvalModK)
        .def_readonly("montgomery", &CKKSParams::montgomery)
        .def_readonly("fold_hybrid_keyswitch", &CKKSParams::fold_hybrid_keyswitch);

    m.def("find_p_and_q_primes", &find_p_and_q_primes);

    py::class_<CKKSContext>(m, "CKKSContext")
        .def(py::init<CKKSParams>())
        .def_static("makeLane", &CKKSContext::makeLane, py::keep_alive<0, 1>())
        .def("isLane", &CKKSContext::isLane)
        .def("getParams", &CKKSContext::getParams, py::return_value_policy::reference_internal)
        .def("getSecretKeyCoeffs", [](const CKKSContext& c) {
            const auto& s = c.getSecretKeyCoeffs();
            py::array_t<int8_t> a(s.size());
            std::copy(s.begin(), s.end(), a.mutable_data());
            return a;
        })
        .def("getSecretKey", [](const CKKSContext& c) {
            std::vector<uint64_t> flat;
            for (auto& l : c.getSecretKey()) flat.insert(flat.end(), l.begin(), l.end());
            return to_np(flat);
        })
        .def("getRelinKey", [](const CKKSContext& c) { return to_np(c.downloadKey(c.getRelinKeyDevice())); })
        .def("getConjugationKey", [](const CKKSContext& c) { return to_np(c.downloadKey(c.getConjugationKeyDevice())); })
        .def("getRotationKey", [](const CKKSContext& c, int64_t r) { return to_np(c.downloadKey(c.getRotationKeyDevice(r))); })
        .def("getRotationKeyValues", &CKKSContext::getRotationKeyValues)
        .def("getRotationKeyCount", &CKKSContext::getRotationKeyCount)
        .def("hasRotationKey", &CKKSContext::hasRotationKey)
        .def("addRotationKey", &CKKSContext::addRotationKey)
        .def("keyDigitCount", &CKKSContext::keyDigitCount)
        .def("getPermTable", [](const CKKSContext& c) { return std::vector<size_t>(c.getPermTable()); })
        .def("setRelinKey", [](CKKSContext& c, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> k) { c.setRelinKey(from_np(k)); })
        .def("setConjugationKey", [](CKKSContext& c, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> k) { c.setConjugationKey(from_np(k)); })
        .def("setRotationKey", [](CKKSContext& c, int64_t r, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> k) { c.setRotationKey(r, from_np(k)); })
        .def("setSecretKeyEval", [](CKKSContext& c, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> k) { c.setSecretKeyEval(from_np(k)); })
        .def("getLowSecretKeyCoeffs", [](const CKKSContext& c) {
            const auto& s = c.getLowSecretKeyCoeffs();
            py::array_t<int8_t> a(s.size());
            std::copy(s.begin(), s.end(), a.mutable_data());
            return a;
        })
        .def("getDenseToSparseSwitchKey", [](const CKKSContext& c) { return to_np(c.downloadDenseToSparseSwitchKey()); })
        .def("getSparseToDenseSwitchKey", [](const CKKSContext& c) { return to_np(c.downloadKey(c.getSparseToDenseSwitchKeyDevice())); })
        .def("setDenseToSparseSwitchKey", [](CKKSContext& c, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> k) { c.setDenseToSparseSwitchKey(from_np(k)); })
        .def("setSparseToDenseSwitchKey", [](CKKSContext& c, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> k) { c.setSparseToDenseSwitchKey(from_np(k)); })
        .def("setEvalModPolynomial", &CKKSContext::setEvalModPolynomial)
        .def("getEvalModS2CPlan", &CKKSContext::getEvalModS2CPlan, py::return_value_policy::reference_internal)
        .def("getEvalModScaledC2SPlan", &CKKSContext::getEvalModScaledC2SPlan, py::return_value_policy::reference_internal)
        .def("getEvalModScaledC2STransformScale", &CKKSContext::getEvalModScaledC2STransformScale)
        .def("tuneReport", [](const CKKSContext& c) {
            std::vector<char> buf(4096);
            mi_check(mi_tune_report(c.handle(), buf.data(), buf.size()), "tuneReport");
            return std::string(buf.data());
        })
        .def("getRotationNttMap", &CKKSContext::getRotationNttMap)
        .def("getConjugationNttMap", &CKKSContext::getConjugationNttMap)
        .def("memoryStats", [](const CKKSContext& c) {
            int64_t st[4];
            mi_check(mi_memory_stats(c.handle(), st), "memoryStats");
            return std::vector<int64_t>(st, st + 4);
        })
        // NTT probe: events around every NTT launch pair between the two calls;
        // the end returns [summed ms, launch pairs, algorithmic bytes, limbs]
        .def("nttProbeBegin", [](const CKKSContext& c) { mi_check(mi_probe_begin(c.handle()), "nttProbeBegin"); })
        .def("nttProbeEnd", [](const CKKSContext& c) {
            double st[4];
            mi_check(mi_probe_end(c.handle(), st), "nttProbeEnd");
            return std::vector<double>(st, st + 4);
        });

    py::class_<CKKSCiphertext>(m, "CKKSCiphertext")
        .def(py::init([](const CKKSContext& c, std::vector<py::array_t<uint64_t, py::array::c_style | py::array::forcecast>> polys,
                         long double scale, size_t level) {
                 MyVector<MyVector<uint64_t>> h;
                 for (auto& p : polys) h.push_back(from_np(p));
                 return CKKSCiphertext(c, std::move(h), scale, level);
             }),
             py::keep_alive<1, 2>())
        .def("clone", &CKKSCiphertext::clone, py::keep_alive<0, 1>())
        .def("getScale", [](const CKKSCiphertext& ct) { return (double)ct.getScale(); })
        .def("getLevel", &CKKSCiphertext::getLevel)
        .def("getN", &CKKSCiphertext::getN)
        .def("getNumPolys", &CKKSCiphertext::getNumPolys)
        .def("getSecretOwner", &CKKSCiphertext::getSecretOwner)
        .def("getMessageEncodingState", &CKKSCiphertext::getMessageEncodingState)
        .def("getPoly", [](const CKKSCiphertext& ct, size_t i) { return to_np(ct.getPoly(i)); })
        .def("getBatchSize", &CKKSCiphertext::getBatchSize)
        .def("getMemberPoly", [](const CKKSCiphertext& ct, size_t i, size_t m) { return to_np(ct.getPoly(i, m)); })
        .def("setPoly", [](CKKSCiphertext& ct, size_t i, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> a) {
            ct.setPoly(i, from_np(a));
        })
        .def("setScale", [](CKKSCiphertext& ct, double s) { ct.setScale(s); });

    py::class_<CKKSEncoding>(m, "CKKSEncoding")
        .def(py::init([](const CKKSContext& c, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> pt,
                         long double scale, size_t level, bool mont) {
                 return CKKSEncoding(c, from_np(pt), scale, level, mont);
             }),
             py::arg("context"), py::arg("plaintext"), py::arg("scale"), py::arg("level"),
             py::arg("montgomery_form") = false, py::keep_alive<1, 2>())
        .def("getPlaintext", [](const CKKSEncoding& e) { return to_np(e.getPlaintext()); })
        .def("getScale", [](const CKKSEncoding& e) { return (double)e.getScale(); })
        .def("getLevel", &CKKSEncoding::getLevel)
        .def("isMontgomeryForm", &CKKSEncoding::isMontgomeryForm);

    m.def("ckks_encode",
          [](const CKKSContext& c, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> v,
             size_t level, double scale, bool eval) {
              return scale > 0 ? ckks_encode(c, complex_vec(v), scale, level, eval)
                               : ckks_encode(c, complex_vec(v), level, eval);
          },
          py::arg("context"), py::arg("values"), py::arg("level"), py::arg("scale") = 0.0, py::arg("eval") = false,
          py::keep_alive<0, 1>());
    m.def("ckks_encode_coefficients",
          [](const CKKSContext& c, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> v,
             double scale, size_t level, bool eval) { return ckks_encode_coefficients(c, complex_vec(v), scale, level, eval); },
          py::arg("context"), py::arg("values"), py::arg("scale"), py::arg("level"), py::arg("eval") = false,
          py::keep_alive<0, 1>());
    m.def("encode_slots_custom_basis",
          [](const CKKSContext& c, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> v,
             double scale, std::vector<uint64_t> moduli) {
              return to_np(encode_slots_custom_basis(c.getParams().getN(), complex_vec(v), scale,
                                                     MyVector<uint64_t>(moduli.begin(), moduli.end()),
                                                     c.getTwiddleFft(), c.getPermTable()));
          });
    m.def("ckks_encrypt",
          [](const CKKSContext& c, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> v,
             size_t level) { return ckks_encrypt(c, complex_vec(v), level); },
          py::keep_alive<0, 1>());
    m.def("ckks_encrypt_plaintext", [](const CKKSContext& c, const CKKSEncoding& e, bool eval) {
              return ckks_encrypt(c, e, eval);
          },
          py::arg("context"), py::arg("encoding"), py::arg("eval") = false, py::keep_alive<0, 1>());
    m.def("ckks_decrypt",
          [](const CKKSContext& c, const CKKSCiphertext& ct, CKKSDecryptDomain d) {
              return complex_np(ckks_decrypt(c, ct, d));
          },
          py::arg("context"), py::arg("ct"), py::arg("domain") = CKKSDecryptDomain::Slots);
    m.def("ckks_decrypt_batch",
          [](const CKKSContext& c, const py::list& cts, CKKSDecryptDomain d) {
              std::vector<const CKKSCiphertext*> ptrs;
              for (const auto& o : cts) ptrs.push_back(&o.cast<const CKKSCiphertext&>());
              std::vector<MyVector<Complex>> out;
              {
                  py::gil_scoped_release unlocked;
                  out = ckks_decrypt_batch(c, ptrs, d);
              }
              py::list res;
              for (auto& v : out) res.append(complex_np(v));
              return res;
          },
          py::arg("context"), py::arg("cts"), py::arg("domain") = CKKSDecryptDomain::Slots);
    // the decrypted coefficients decoded on the host (the reference's
    // arithmetic, the ciphertext's exact scale): the check of the device decode
    m.def("ckks_decrypt_host_decode",
          [](const CKKSContext& c, const CKKSCiphertext& ct, CKKSDecryptDomain d) {
              return complex_np(ckks_decode_plaintext(c, ckks_decrypt_coefficients(c, ct), ct.getLevel() + 1,
                                                      ct.getScale(), d));
          },
          py::arg("context"), py::arg("ct"), py::arg("domain") = CKKSDecryptDomain::Slots);
    m.def("ckks_decode_plaintext",
          [](const CKKSContext& c, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> coeffs, size_t limbs,
             double scale, CKKSDecryptDomain d) {
              return complex_np(ckks_decode_plaintext(c, from_np(coeffs), limbs, scale, d));
          },
          py::arg("context"), py::arg("coeffs"), py::arg("limb_count"), py::arg("scale"),
          py::arg("domain") = CKKSDecryptDomain::Slots);

    m.def("twiddle_ifft_generation", [](size_t N) { return complex_np(twiddle_ifft_generation(N)); });
    m.def("perm_table_generation", [](size_t N) { return std::vector<size_t>(perm_table_generation(N)); });
    m.def("encode_slots_custom_basis_tables",
          [](size_t N, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> v, double scale,
             std::vector<uint64_t> moduli) {
              auto tw = twiddle_ifft_generation(N);
              for (auto& z : tw) z = std::conj(z);
              return to_np(encode_slots_custom_basis(N, complex_vec(v), scale,
                                                     MyVector<uint64_t>(moduli.begin(), moduli.end()), tw,
                                                     perm_table_generation(N)));
          });
    m.def("decode_plaintext_custom_basis",
          [](size_t N, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> coeffs,
             std::vector<uint64_t> moduli, double scale, CKKSDecryptDomain d) {
              return complex_np(decode_plaintext_custom_basis(N, from_np(coeffs),
                                                              MyVector<uint64_t>(moduli.begin(), moduli.end()), scale,
                                                              twiddle_ifft_generation(N), perm_table_generation(N), d));
          },
          py::arg("N"), py::arg("coeffs"), py::arg("moduli"), py::arg("scale"),
          py::arg("domain") = CKKSDecryptDomain::Slots);
    m.def("ckks_mul_inplace", py::overload_cast<const CKKSContext&, CKKSCiphertext&, const CKKSCiphertext&>(&ckks_mul_inplace));
    m.def("ckks_mul_plain_inplace", py::overload_cast<const CKKSContext&, CKKSCiphertext&, const CKKSEncoding&>(&ckks_mul_inplace));
    m.def("ckks_mul_int_inplace", py::overload_cast<const CKKSContext&, CKKSCiphertext&, int64_t>(&ckks_mul_inplace));
    m.def("ckks_mul_const_inplace", py::overload_cast<const CKKSContext&, CKKSCiphertext&, Complex, double>(&ckks_mul_inplace));
    m.def("ckks_square_inplace", &ckks_square_inplace);
    m.def("ckks_rescale", &ckks_rescale);
    m.def("ckks_drop_to_level_inplace", &ckks_drop_to_level_inplace);
    m.def("ckks_add_inplace", py::overload_cast<const CKKSContext&, CKKSCiphertext&, const CKKSCiphertext&>(&ckks_add_inplace));
    m.def("ckks_add_plain_inplace", py::overload_cast<const CKKSContext&, CKKSCiphertext&, const CKKSEncoding&>(&ckks_add_inplace));
    m.def("ckks_sub_inplace", &ckks_sub_inplace);
    m.def("ckks_add_integer_constant_inplace", &ckks_add_integer_constant_inplace);
    m.def("ckks_relin_hybrid_inplace", &ckks_relin_hybrid_inplace);
    m.def("ckks_rotate_inplace", &ckks_rotate_inplace);
    m.def("ckks_rotate", &ckks_rotate, py::keep_alive<0, 1>());
    m.def("ckks_conjugate_inplace", &ckks_conjugate_inplace);
    m.def("ckks_conjugate", &ckks_conjugate, py::keep_alive<0, 1>());
    py::enum_<CKKSBootstrapDftType>(m, "CKKSBootstrapDftType")
        .value("SlotsToCoeffs", CKKSBootstrapDftType::SlotsToCoeffs)
        .value("CoeffsToSlots", CKKSBootstrapDftType::CoeffsToSlots);
    py::class_<CKKSBootstrapDftMatrixLiteral>(m, "CKKSBootstrapDftMatrixLiteral")
        .def(py::init<>())
        .def_readwrite("type", &CKKSBootstrapDftMatrixLiteral::type)
        .def_readwrite("level", &CKKSBootstrapDftMatrixLiteral::level)
        .def_readwrite("baby_step_count", &CKKSBootstrapDftMatrixLiteral::baby_step_count)
        .def_readwrite("stage_count", &CKKSBootstrapDftMatrixLiteral::stage_count)
        .def_readwrite("log_plaintext_scale_groups", &CKKSBootstrapDftMatrixLiteral::log_plaintext_scale_groups)
        .def_readwrite("merge_depths", &CKKSBootstrapDftMatrixLiteral::merge_depths)
        .def_readwrite("log_bsgs_ratio", &CKKSBootstrapDftMatrixLiteral::log_bsgs_ratio)
        .def_readwrite("active_slots", &CKKSBootstrapDftMatrixLiteral::active_slots)
        .def_readwrite("transform_scale", &CKKSBootstrapDftMatrixLiteral::transform_scale)
        .def_readwrite("plaintext_scale", &CKKSBootstrapDftMatrixLiteral::plaintext_scale)
        .def_readwrite("input_selector_start", &CKKSBootstrapDftMatrixLiteral::input_selector_start)
        .def_readwrite("input_selector_size", &CKKSBootstrapDftMatrixLiteral::input_selector_size)
        .def_readwrite("rescale_after_each_stage", &CKKSBootstrapDftMatrixLiteral::rescale_after_each_stage);
    py::class_<CKKSLinearTransformOwnedStage>(m, "CKKSLinearTransformOwnedStage")
        .def_readonly("diagonal_indices", &CKKSLinearTransformOwnedStage::diagonal_indices)
        .def_readonly("slot_count", &CKKSLinearTransformOwnedStage::slot_count)
        .def_readonly("baby_step_count", &CKKSLinearTransformOwnedStage::baby_step_count)
        .def_readonly("qp_limb_count", &CKKSLinearTransformOwnedStage::qp_limb_count)
        .def_readonly("level", &CKKSLinearTransformOwnedStage::level)
        .def_readonly("plaintext_scale", &CKKSLinearTransformOwnedStage::plaintext_scale)
        .def_readonly("rescale_after", &CKKSLinearTransformOwnedStage::rescale_after)
        .def_readonly("output_encoding_state", &CKKSLinearTransformOwnedStage::output_encoding_state)
        .def("downloadDiagonalQp", [](const CKKSLinearTransformOwnedStage& st, const CKKSContext& c, size_t t) {
            return to_np(st.downloadDiagonalQp(c, t));
        });
    py::class_<CKKSLinearTransformPlan>(m, "CKKSLinearTransformPlan", py::module_local(false))
        .def_readwrite("use_qp_evaluator", &CKKSLinearTransformPlan::use_qp_evaluator)
        .def("stageCount", [](const CKKSLinearTransformPlan& p) { return p.stages.size(); })
        .def("stage", [](const CKKSLinearTransformPlan& p, size_t i) -> const CKKSLinearTransformOwnedStage& {
                 return p.stages.at(i);
             },
             py::return_value_policy::reference_internal);
    auto as_bytes = [](const std::vector<uint8_t>& v) {
        return py::bytes(reinterpret_cast<const char*>(v.data()), v.size());
    };
    auto span_of = [](const py::bytes& b) {
        std::string_view sv(b);
        return std::span<const uint8_t>(reinterpret_cast<const uint8_t*>(sv.data()), sv.size());
    };
    m.def("ckks_serialize_ciphertext", [as_bytes](const CKKSCiphertext& ct) { return as_bytes(ckks_serialize_ciphertext(ct)); });
    m.def("ckks_deserialize_ciphertext",
          [span_of](const CKKSContext& c, const py::bytes& b) { return ckks_deserialize_ciphertext(c, span_of(b)); },
          py::keep_alive<0, 1>());
    m.def("ckks_serialize_encoding", [as_bytes](const CKKSEncoding& pt) { return as_bytes(ckks_serialize_encoding(pt)); });
    m.def("ckks_deserialize_encoding",
          [span_of](const CKKSContext& c, const py::bytes& b) { return ckks_deserialize_encoding(c, span_of(b)); },
          py::keep_alive<0, 1>());
    m.def("ckks_serialize_evaluation_keys",
          [as_bytes](const CKKSContext& c) { return as_bytes(ckks_serialize_evaluation_keys(c)); });
    m.def("ckks_load_evaluation_keys",
          [span_of](CKKSContext& c, const py::bytes& b) { ckks_load_evaluation_keys(c, span_of(b)); });
    m.def("ckks_generate_bootstrap_dft_plan", &ckks_generate_bootstrap_dft_plan, py::keep_alive<0, 1>());
    m.def("ckks_encode_linear_transform_plan",
          [](const CKKSContext& c, const std::map<size_t, std::vector<std::complex<double>>>& diags, size_t level,
             size_t baby_step_count, double plaintext_scale, bool use_qp, bool rescale_after) {
              std::map<size_t, MyVector<Complex>> m;
              for (const auto& [i, v] : diags) m[i] = MyVector<Complex>(v.begin(), v.end());
              return ckks_encode_linear_transform_plan(c, m, level, baby_step_count, plaintext_scale, use_qp,
                                                       rescale_after);
          },
          py::arg("context"), py::arg("diagonals"), py::arg("level"), py::arg("baby_step_count"),
          py::arg("plaintext_scale"), py::arg("use_qp_evaluator") = true, py::arg("rescale_after") = true,
          py::keep_alive<0, 1>());
    m.def("ckks_apply_linear_transform_plan",
          [](const CKKSContext& c, const CKKSCiphertext& ct, const CKKSLinearTransformPlan& plan) {
              return ckks_apply_linear_transform_plan(c, ct.clone(), plan);
          },
          py::keep_alive<0, 1>());
    m.def("ckks_apply_linear_transform_plan_stage",
          [](const CKKSContext& c, const CKKSCiphertext& ct, const CKKSLinearTransformPlan& plan, size_t i) {
              const auto& st = plan.stages.at(i);
              return plan.use_qp_evaluator ? ckks_apply_linear_transform_owned_stage_qp(c, ct, st)
                                           : ckks_apply_linear_transform_owned_stage_q(c, ct, st);
          },
          py::keep_alive<0, 1>());
    // ckks_apply_linear_transform_bsgs over diagonals given as plaintext limbs
    // (each becomes a CKKSEncoding at the ciphertext's level)
    m.def("ckks_apply_linear_transform_bsgs",
          [](const CKKSContext& c, const CKKSCiphertext& ct, const std::vector<py::array_t<uint64_t>>& pts,
             double scale, bool montgomery, const std::vector<size_t>& indices, size_t baby_step_count,
             size_t slot_count, bool rescale_after) {
              MyVector<CKKSEncoding> diags;
              diags.reserve(pts.size());
              for (const auto& a : pts) {
                  auto arr = py::array_t<uint64_t, py::array::c_style | py::array::forcecast>::ensure(a);
                  diags.emplace_back(c, from_np(arr), (long double)scale, ct.getLevel(), montgomery);
              }
              CKKSLinearTransformStage st;
              st.diagonals = std::span<const CKKSEncoding>(diags.data(), diags.size());
              st.diagonal_indices = std::span<const size_t>(indices.data(), indices.size());
              st.slot_count = slot_count;
              st.baby_step_count = baby_step_count;
              st.rescale_after = rescale_after;
              return ckks_apply_linear_transform_bsgs(c, ct, st);
          },
          py::arg("context"), py::arg("ct"), py::arg("diagonals"), py::arg("scale"), py::arg("montgomery"),
          py::arg("diagonal_indices"), py::arg("baby_step_count"), py::arg("slot_count") = 0,
          py::arg("rescale_after") = true, py::keep_alive<0, 1>());
    m.def("ckks_key_switch_dense_to_sparse_inplace", &ckks_key_switch_dense_to_sparse_inplace);
    py::class_<CKKSGraph>(m, "CKKSGraph")
        .def("launch", &CKKSGraph::launch)
        .def("reset", &CKKSGraph::reset);
    m.def("ckks_capture_graph",
          [](const CKKSContext& c, py::function body) { return ckks_capture_graph(c, [&] { body(); }); },
          py::keep_alive<0, 1>());
    m.def("ckks_copy_into", &ckks_copy_into);
    // the keyswitch-level API: keys by name (rotation r, "relin", "conj"),
    // polynomials as numpy arrays (copied in, results returned)
    auto key_of = [](const CKKSContext& c, py::object which) -> CKKSSwitchingKey {
        if (py::isinstance<py::str>(which)) {
            const std::string w = which.cast<std::string>();
            if (w == "relin") return c.getRelinHybrid(0);
            if (w == "conj") return c.getConjugationKey();
            if (w == "d2s") return c.getDenseToSparseSwitchKey();
            if (w == "s2d") return c.getSparseToDenseSwitchKey();
            throw std::invalid_argument("unknown switching key " + w);
        }
        return c.getRotationKey(which.cast<int64_t>());
    };
    auto combine = [](int m) { return m ? CKKSHybridKeySwitchCombine::Assign : CKKSHybridKeySwitchCombine::Add; };
    m.def("ckks_hybrid_key_switch_inplace",
          [key_of, combine](const CKKSContext& c, size_t level, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> kt,
                            py::object key, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> b,
                            py::array_t<uint64_t, py::array::c_style | py::array::forcecast> a, int bm, int am, bool mont) {
              MyVector<uint64_t> k = from_np(kt), bo = from_np(b), ao = from_np(a);
              ckks_hybrid_key_switch_inplace(c, level, k, key_of(c, key), bo, ao, combine(bm), combine(am), mont);
              return py::make_tuple(to_np(bo), to_np(ao));
          });
    m.def("ckks_key_switch_inplace", [key_of](const CKKSContext& c, CKKSCiphertext& ct, py::object key, bool mont) {
        ckks_key_switch_inplace(c, ct, key_of(c, key), mont);
    });
    py::class_<CKKSEvalHoistedCiphertextQP>(m, "CKKSEvalHoistedCiphertextQP")
        .def(py::init<>())
        .def_readonly("level", &CKKSEvalHoistedCiphertextQP::level)
        .def_readonly("q_limb_count", &CKKSEvalHoistedCiphertextQP::q_limb_count)
        .def_readonly("p_limb_count", &CKKSEvalHoistedCiphertextQP::p_limb_count);
    m.def("ckks_hoist_ciphertext_a_qp_eval", [](const CKKSContext& c, const CKKSCiphertext& ct) {
        CKKSHoistWorkspace ws;
        CKKSEvalHoistedCiphertextQP out;
        ckks_hoist_ciphertext_a_qp_eval(c, ct, ws, out);
        return out;
    }, py::keep_alive<0, 1>());
    m.def("ckks_hybrid_key_switch_hoisted_inplace",
          [key_of, combine](const CKKSContext& c, const CKKSEvalHoistedCiphertextQP& h, py::object key,
                            py::array_t<uint64_t, py::array::c_style | py::array::forcecast> b,
                            py::array_t<uint64_t, py::array::c_style | py::array::forcecast> a, int bm, int am, bool mont,
                            std::vector<size_t> ntt_map) {
              MyVector<uint64_t> bo = from_np(b), ao = from_np(a);
              ckks_hybrid_key_switch_hoisted_inplace(c, h, key_of(c, key), bo, ao, combine(bm), combine(am), mont,
                                                     std::span<const size_t>(ntt_map));
              return py::make_tuple(to_np(bo), to_np(ao));
          });
    m.def("ckks_hybrid_key_switch_qp_from_key_term",
          [key_of](const CKKSContext& c, size_t level, py::array_t<uint64_t, py::array::c_style | py::array::forcecast> kt,
                   py::object key, bool mont) {
              MyVector<uint64_t> b, a;
              ckks_hybrid_key_switch_qp_from_key_term(c, level, from_np(kt), key_of(c, key), b, a, mont);
              return py::make_tuple(to_np(b), to_np(a));
          });
    m.def("ckks_apply_ntt_automorphism_to_flat_rns_poly_inplace",
          [](py::array_t<uint64_t, py::array::c_style | py::array::forcecast> poly, size_t N, size_t limbs,
             std::vector<size_t> map) {
              MyVector<uint64_t> p = from_np(poly);
              ckks_apply_ntt_automorphism_to_flat_rns_poly_inplace(p, N, limbs, std::span<const size_t>(map));
              return to_np(p);
          });
    m.def("ckks_key_switch_sparse_to_dense_inplace", &ckks_key_switch_sparse_to_dense_inplace);
    m.def("ckks_modraise_inplace", &ckks_modraise_inplace);
    m.def("ckks_s2c_generalized", &ckks_s2c_generalized, py::keep_alive<0, 1>());
    m.def("ckks_c2s_generalized", &ckks_c2s_generalized, py::keep_alive<0, 1>());
    m.def("ckks_evalmod_binary_inplace", &ckks_evalmod_binary_inplace);
    // a pair of ciphertexts, each keeping the context alive (a tuple itself
    // cannot carry keep_alive)
    auto ct_pair = [](std::pair<CKKSCiphertext, CKKSCiphertext> v, const py::object& context) {
        py::object a = py::cast(std::move(v.first)), b = py::cast(std::move(v.second));
        py::detail::keep_alive_impl(a, context);
        py::detail::keep_alive_impl(b, context);
        return py::make_tuple(a, b);
    };
    m.def("ckks_bootstrap_pre_evalmod", [ct_pair](const py::object& context, const CKKSCiphertext& ct) {
        return ct_pair(ckks_bootstrap_pre_evalmod(context.cast<const CKKSContext&>(), ct.clone()), context);
    });
    m.def("ckks_batch_bootstrap_evalmod_from_complex_packed",
          [ct_pair](const py::object& context, const CKKSCiphertext& ct) {
              return ct_pair(ckks_batch_bootstrap_evalmod_from_complex_packed(context.cast<const CKKSContext&>(),
                                                                              ct.clone()),
                             context);
          });
    m.def("ckks_batch_bootstrap_evalmod_inplace", &ckks_batch_bootstrap_evalmod_inplace);
    // Walsh S-box helpers (lists of ciphertexts are cloned into the call)
    auto ct_list = [](const py::list& l) {
        MyVector<CKKSCiphertext> v;
        v.reserve(l.size());
        for (const auto& h : l) v.push_back(h.cast<const CKKSCiphertext&>().clone());
        return v;
    };
    m.def("ckks_aes_walsh_table", []() { return ckks_aes_walsh_table(); });
    m.def("aes_scale_up_factor", [](double from, double to) {
        int64_t factor = 0;
        const char* err = ckks_detail::aes_scale_up_factor(from, to, &factor);
        return py::make_tuple(factor, err ? py::cast(std::string(err)) : py::none());
    });
    m.def("ckks_walsh_sum_for_output_pair",
          [ct_list](const CKKSContext& c, const py::list& low, const py::list& high, size_t real_bit, size_t imag_bit,
                    const CKKSEncoding& encoded_i) {
              return ckks_walsh_sum_for_output_pair(c, ct_list(low), ct_list(high), real_bit, imag_bit, encoded_i);
          },
          py::keep_alive<0, 1>());
    m.def("ckks_walsh_sum_to_bit_pair",
          [](const CKKSContext& c, const CKKSCiphertext& ct) { return ckks_walsh_sum_to_bit_pair(c, ct.clone()); },
          py::keep_alive<0, 1>());
    // each returned ciphertext keeps the context alive (its device blocks
    // are freed through it), as keep_alive<0, 1> does for single results
    auto ct_pylist = [](MyVector<CKKSCiphertext>& v, const py::object& context) {
        py::list r;
        for (auto& x : v) {
            py::object o = py::cast(std::move(x));
            py::detail::keep_alive_impl(o, context);
            r.append(o);
        }
        return r;
    };
    // a batch of copies of the listed ciphertexts, and copies of a batch's members
    m.def("ckks_stack", [ct_list](const py::object& context, const py::list& cts) {
        py::object o = py::cast(CKKSCiphertext::stack(ct_list(cts)));
        py::detail::keep_alive_impl(o, context);
        return o;
    });
    m.def("ckks_unstack", [ct_pylist](const py::object& context, const CKKSCiphertext& ct) {
        auto v = ct.clone().unstack();
        return ct_pylist(v, context);
    });
    m.def("ckks_conjugate_batch", [ct_list, ct_pylist](const py::object& context, const py::list& cts) {
        auto out = ckks_conjugate_batch(context.cast<const CKKSContext&>(), ct_list(cts));
        return ct_pylist(out, context);
    });
    m.def("ckks_walsh_recombine_signs_to_paired_outputs",
          [ct_list, ct_pylist](const py::object& context, const py::list& low, const py::list& high) {
              const CKKSContext& c = context.cast<const CKKSContext&>();
              auto out = ckks_walsh_recombine_signs_to_paired_outputs(c, ct_list(low), ct_list(high));
              return ct_pylist(out, context);
          });
    // in-place functions over a list: the ciphertexts are moved into a
    // vector for the call and moved back into their Python objects
    auto inplace_list = [](const py::list& l, auto&& fn) {
        MyVector<CKKSCiphertext> v;
        v.reserve(l.size());
        for (const auto& h : l) v.push_back(std::move(h.cast<CKKSCiphertext&>()));
        auto put_back = [&]() {
            size_t i = 0;
            for (const auto& h : l) h.cast<CKKSCiphertext&>() = std::move(v[i++]);
        };
        try {
            fn(v);
        } catch (...) {
            put_back();
            throw;
        }
        put_back();
    };
    m.def("ckks_aes_mask_inactive_slots_inplace",
          [inplace_list](const CKKSContext& c, const py::list& state, size_t block_count) {
              inplace_list(state, [&](MyVector<CKKSCiphertext>& v) {
                  ckks_aes_mask_inactive_slots_inplace(c, v, block_count);
              });
          });
    m.def("ckks_aes_mask_final_refresh_slots_inplace",
          [inplace_list](const CKKSContext& c, const py::list& bits, size_t active, size_t minimum_level_to_keep) {
              inplace_list(bits, [&](MyVector<CKKSCiphertext>& v) {
                  ckks_aes_mask_final_refresh_slots_inplace(c, v, active, minimum_level_to_keep);
              });
          });
    m.def("ckks_walsh_sbox_complex_blocks_mask_packed",
          [ct_list, ct_pylist](const py::object& context, const py::list& input, size_t block_count) {
              CKKSContext& c = context.cast<CKKSContext&>();
              auto out = ckks_walsh_sbox_complex_blocks_mask_packed(c, ct_list(input), block_count);
              return ct_pylist(out, context);
          });
    m.def("ckks_walsh_build_mask_packed_parity_input",
          [ct_list](const CKKSContext& c, const py::list& input, size_t block_count) {
              return ckks_walsh_build_mask_packed_parity_input(c, ct_list(input), block_count);
          },
          py::keep_alive<0, 1>());
    m.def("ckks_walsh_unpack_mask_packed_signs",
          [ct_pylist](const py::object& context, const CKKSCiphertext& packed, size_t block_count) {
              const CKKSContext& c = context.cast<const CKKSContext&>();
              auto [low, high] = ckks_walsh_unpack_mask_packed_signs(c, packed, block_count);
              return py::make_tuple(ct_pylist(low, context), ct_pylist(high, context));
          });
    m.def("rotation_automorphism_index", &rotation_automorphism_index);
    m.def("ckks_evalchebyshev_preconditioned",
          [](const CKKSContext& c, const CKKSCiphertext& x, std::vector<double> coeffs, double target_scale) {
              CKKSEvalPolyOptions o;
              o.target_scale = target_scale;
              return ckks_evalchebyshev_preconditioned(c, x, coeffs, o);
          },
          py::arg("context"), py::arg("x"), py::arg("coefficients"), py::arg("target_scale") = 0.0,
          py::keep_alive<0, 1>());

    py::class_<AesRoundKeyBits>(m, "CKKSAesRoundKeyBits")
        .def("__len__", [](const AesRoundKeyBits& k) { return k.bits.size(); })
        .def("get", [](const AesRoundKeyBits& k, size_t i) { return k.bits.at(i).clone(); }, py::keep_alive<0, 1>());
    m.def("aes128_expand_key", &aes128_expand_key);
    m.def("aes128_encrypt_block", &aes128_encrypt_block);
    m.def("aes128_encrypt_block_expanded", &aes128_encrypt_block_expanded);
    m.def("aes128_ctr_crypt_block", &aes128_ctr_crypt_block);
    m.def("default_evalmod_polynomial", &default_evalmod_polynomial);
    m.def("evalmod_polynomial_is_reference_fit", &evalmod_polynomial_is_reference_fit);
    m.def("aes_sbox_table", []() { return aes_sbox_table(); });
    m.def("ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks",
          [](const CKKSContext& c, const std::array<uint8_t, 176>& rk, size_t active) {
              return AesRoundKeyBits(ckks_aes128_encrypt_expanded_round_key_bits_walsh_complex_blocks(c, rk, active));
          },
          py::keep_alive<0, 1>());
    py::class_<CKKSAesCtrServer>(m, "CKKSAesCtrServer")
        .def(py::init([](CKKSContext& owner, const std::array<uint8_t, 16>& key, size_t lanes, int64_t budget,
                         size_t max_request_blocks, size_t max_queued_evaluations) {
                 CKKSAesServerOptions o;
                 o.lanes = lanes;
                 o.memory_budget_bytes = budget;
                 o.max_request_blocks = max_request_blocks;
                 o.max_queued_evaluations = max_queued_evaluations;
                 py::gil_scoped_release unlocked;
                 return std::make_unique<CKKSAesCtrServer>(owner, key, o);
             }),
             py::arg("owner"), py::arg("aes_key"), py::arg("lanes") = 2, py::arg("memory_budget_bytes") = 0,
             py::arg("max_request_blocks") = 1024, py::arg("max_queued_evaluations") = 4, py::keep_alive<1, 2>())
        .def("submit",
             [](CKKSAesCtrServer& s, const py::array_t<uint8_t, py::array::c_style>& counters,
                const py::array_t<uint8_t, py::array::c_style>& cts) {
                 auto c = blocks_from_array(counters), x = blocks_from_array(cts);
                 return s.submit(std::move(c), std::move(x));
             })
        .def("wait",
             [](CKKSAesCtrServer& s, uint64_t id) {
                 CKKSAesRequestResult r;
                 {
                     py::gil_scoped_release unlocked;
                     r = s.wait(id);
                 }
                 py::dict d;
                 d["id"] = r.id;
                 d["refused"] = r.refused;
                 d["error"] = r.error;
                 const size_t n = r.plaintexts.size();
                 py::array_t<uint8_t> pt({(py::ssize_t)n, (py::ssize_t)16});
                 py::array_t<double> vals({(py::ssize_t)n, (py::ssize_t)128});
                 for (size_t j = 0; j < n; ++j) {
                     std::memcpy(pt.mutable_data(j, 0), r.plaintexts[j].data(), 16);
                     std::memcpy(vals.mutable_data(j, 0), r.values[j].data(), 128 * sizeof(double));
                 }
                 d["plaintexts"] = pt;
                 d["values"] = vals;
                 d["arrival"] = r.arrival;
                 d["start"] = r.start;
                 d["finish"] = r.finish;
                 d["lane"] = r.lane;
                 return d;
             })
        .def("laneCount", &CKKSAesCtrServer::laneCount)
        .def("evaluationBytes", &CKKSAesCtrServer::evaluationBytes)
        .def("evaluations", &CKKSAesCtrServer::evaluations);
    m.def("ckks_aes_serve_take", &ckks_aes_serve_take);
    m.def("steady_clock_seconds", [] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    });
    m.def("ckks_aes128_encrypt_expanded_round_key_bits",
          [](const CKKSContext& c, const std::array<uint8_t, 176>& rk, size_t first_level) {
              AesRoundKeyBits k(MyVector<CKKSCiphertext>{});
              {
                  py::gil_scoped_release unlocked;
                  k.bits = ckks_aes128_encrypt_expanded_round_key_bits(c, rk, first_level);
              }
              return k;
          },
          py::keep_alive<0, 1>());
    m.def("ckks_aes_ctr",
          [ct_pylist](const py::object& context, const AesRoundKeyBits& keys,
                      const py::array_t<uint8_t, py::array::c_style>& counters,
                      const py::array_t<uint8_t, py::array::c_style>& ciphertexts, bool batch) {
              CKKSContext& c = context.cast<CKKSContext&>();
              const auto ctr = blocks_from_array(counters), cts = blocks_from_array(ciphertexts);
              CKKSAes128CtrOptions options;
              options.batch_independent_work = batch;
              MyVector<CKKSCiphertext> out;
              {
                  py::gil_scoped_release unlocked;
                  out = ckks_aes_ctr(c, keys.bits, ctr, cts, options);
              }
              return ct_pylist(out, context);
          },
          py::arg("context"), py::arg("keys"), py::arg("counter_blocks"), py::arg("ciphertext_blocks"),
          py::arg("batch_independent_work") = true);
    m.def("ckks_aes_ctr_walsh_sbox_complex_blocks",
          [ct_pylist](const py::object& context, const AesRoundKeyBits& keys,
                      const py::array_t<uint8_t, py::array::c_style>& counters,
                      const py::array_t<uint8_t, py::array::c_style>& ciphertexts, const py::list& sbox_contexts,
                      bool fuse_linear_layer, bool batch_sboxes) {
              CKKSContext& c = context.cast<CKKSContext&>();
              const auto ctr = blocks_from_array(counters), cts = blocks_from_array(ciphertexts);
              CKKSAesCtrWalshSboxOptions options;
              for (const auto& w : sbox_contexts) options.sbox_contexts.push_back(&w.cast<CKKSContext&>());
              options.fuse_linear_layer = fuse_linear_layer;
              options.batch_sboxes = batch_sboxes;
              MyVector<CKKSCiphertext> out;
              {
                  py::gil_scoped_release unlocked;
                  out = ckks_aes_ctr_walsh_sbox_complex_blocks(c, keys.bits, ctr, cts, options);
              }
              return ct_pylist(out, context);
          },
          py::arg("context"), py::arg("keys"), py::arg("counter_blocks"), py::arg("ciphertext_blocks"),
          py::arg("sbox_contexts") = py::list(), py::arg("fuse_linear_layer") = true,
          py::arg("batch_sboxes") = true);
}
