// Byte formats for ciphertexts, plaintexts and evaluation keys.
//
// The reference keeps everything in process memory and has no wire or disk
// format; these let fixtures, serving restarts and other processes reuse
// device-resident material.  Layout (little-endian, 8-byte aligned fields):
//
//   magic[8] | u32 version | u32 kind-specific flags | u64 context fingerprint
//   | payload | u64 FNV-1a of every preceding byte
//
// The fingerprint is FNV-1a over N, the Q moduli and the P moduli, so a blob
// loads only into a context with the same ring and chain.  Residues are
// written limb-major exactly as they sit in HBM; scales are stored as the
// 80-bit x87 long double in a zero-padded 16-byte field.
#include <cstring>

#include "ckks.hpp"

namespace {

constexpr uint32_t kVersion = 1;
constexpr char kMagicCt[8] = {'M', 'I', 'C', 'K', 'K', 'S', 'C', 'T'};
constexpr char kMagicPt[8] = {'M', 'I', 'C', 'K', 'K', 'S', 'P', 'T'};
constexpr char kMagicKeys[8] = {'M', 'I', 'C', 'K', 'K', 'S', 'E', 'K'};

enum KeyKind : uint32_t { kRelin = 0, kConj = 1, kRotation = 2, kDenseToSparse = 3, kSparseToDense = 4 };

uint64_t fnv1a(const uint8_t* p, size_t n, uint64_t h = 0xcbf29ce484222325ULL) {
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

uint64_t fingerprint(const CKKSParams& P) {
    std::vector<uint64_t> v;
    v.push_back(P.getN());
    v.push_back(P.getModuli().size());
    v.insert(v.end(), P.getModuli().begin(), P.getModuli().end());
    v.push_back(P.getPModuli().size());
    v.insert(v.end(), P.getPModuli().begin(), P.getPModuli().end());
    return fnv1a(reinterpret_cast<const uint8_t*>(v.data()), v.size() * 8);
}

class Writer {
public:
    explicit Writer(const char (&magic)[8], uint32_t flags, uint64_t fp) {
        bytes_.insert(bytes_.end(), magic, magic + 8);
        u32(kVersion);
        u32(flags);
        u64(fp);
    }
    void u32(uint32_t v) { raw(&v, 4); }
    void u64(uint64_t v) { raw(&v, 8); }
    void i64(int64_t v) { raw(&v, 8); }
    void scale(long double s) {
        uint8_t b[16] = {0};
        std::memcpy(b, &s, 10);  // x87 extended precision: 10 significant bytes
        raw(b, 16);
    }
    void words(const MyVector<uint64_t>& w) { raw(w.data(), w.size() * 8); }
    std::vector<uint8_t> finish() {
        u64(fnv1a(bytes_.data(), bytes_.size()));
        return std::move(bytes_);
    }

private:
    void raw(const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        bytes_.insert(bytes_.end(), b, b + n);
    }
    std::vector<uint8_t> bytes_;
};

class Reader {
public:
    Reader(std::span<const uint8_t> b, const char (&magic)[8], uint64_t fp, const char* what) : b_(b), what_(what) {
        if (b.size() < 32) fail("truncated blob");
        end_ = b.size() - 8;
        const uint64_t stored = load<uint64_t>(end_);
        if (fnv1a(b.data(), end_) != stored) fail("checksum mismatch");
        if (std::memcmp(b.data(), magic, 8) != 0) fail("wrong magic (not this kind of blob)");
        pos_ = 8;
        if (u32() != kVersion) fail("unsupported format version");
        flags_ = u32();
        if (u64() != fp) fail("blob was written for a different ring or modulus chain");
    }
    uint32_t flags() const { return flags_; }
    uint32_t u32() { return next<uint32_t>(); }
    uint64_t u64() { return next<uint64_t>(); }
    int64_t i64() { return next<int64_t>(); }
    long double scale() {
        need(16);
        long double s = 0.0L;
        std::memcpy(&s, b_.data() + pos_, 10);
        pos_ += 16;
        return s;
    }
    MyVector<uint64_t> words(size_t n) {
        if (n > (end_ - pos_) / 8) fail("truncated payload");
        MyVector<uint64_t> w(n);
        std::memcpy(w.data(), b_.data() + pos_, n * 8);
        pos_ += n * 8;
        return w;
    }
    void done() const {
        if (pos_ != end_) fail("trailing bytes after the payload");
    }
    [[noreturn]] void fail(const std::string& m) const { throw std::invalid_argument(std::string(what_) + ": " + m); }

private:
    template <class T>
    T load(size_t at) const {
        T v;
        std::memcpy(&v, b_.data() + at, sizeof(T));
        return v;
    }
    void need(size_t n) const {
        if (pos_ + n > end_) fail("truncated payload");
    }
    template <class T>
    T next() {
        need(sizeof(T));
        T v = load<T>(pos_);
        pos_ += sizeof(T);
        return v;
    }
    std::span<const uint8_t> b_;
    const char* what_;
    size_t pos_ = 0, end_ = 0;
    uint32_t flags_ = 0;
};

void check_residues(const MyVector<uint64_t>& w, const MyVector<uint64_t>& moduli, size_t N, size_t limbs,
                    const Reader& r) {
    for (size_t l = 0; l < limbs; ++l)
        for (size_t i = 0; i < N; ++i)
            if (w[l * N + i] >= moduli[l]) r.fail("residue not below its modulus");
}

}  // namespace

std::vector<uint8_t> ckks_serialize_ciphertext(const CKKSCiphertext& ct) {
    const auto& P = ct.getContext().getParams();
    Writer w(kMagicCt, 0, fingerprint(P));
    w.u64(ct.getN());
    w.u64(ct.getLevel());
    w.u64(ct.getNumPolys());
    w.u64((uint64_t)ct.getSecretOwner() | ((uint64_t)ct.getMessageEncodingState() << 8));
    w.scale(ct.getScale());
    for (size_t k = 0; k < ct.getNumPolys(); ++k) w.words(ct.getPoly(k));
    return w.finish();
}

CKKSCiphertext ckks_deserialize_ciphertext(const CKKSContext& c, std::span<const uint8_t> bytes) {
    const auto& P = c.getParams();
    Reader r(bytes, kMagicCt, fingerprint(P), "ckks_deserialize_ciphertext");
    const uint64_t N = r.u64(), level = r.u64(), npolys = r.u64(), meta = r.u64();
    if (N != P.getN()) r.fail("ring dimension mismatch");
    if (level > P.getMaxLevel()) r.fail("level exceeds the modulus chain");
    if (npolys < 2 || npolys > 3) r.fail("a ciphertext has 2 or 3 polynomials");
    const uint64_t owner = meta & 0xff, state = (meta >> 8) & 0xff;
    if (owner > 1 || state > 1 || (meta >> 16) != 0) r.fail("invalid metadata");
    const long double scale = r.scale();
    MyVector<MyVector<uint64_t>> polys;
    for (uint64_t k = 0; k < npolys; ++k) {
        polys.push_back(r.words((level + 1) * N));
        check_residues(polys.back(), P.getModuli(), N, level + 1, r);
    }
    r.done();
    return CKKSCiphertext(c, std::move(polys), scale, level, (CKKSSecretOwner)owner, (CKKSMessageEncodingState)state);
}

std::vector<uint8_t> ckks_serialize_encoding(const CKKSEncoding& pt) {
    Writer w(kMagicPt, pt.isMontgomeryForm() ? 1u : 0u, fingerprint(pt.getContext().getParams()));
    w.u64(pt.getN());
    w.u64(pt.getLevel());
    w.scale(pt.getScale());
    w.words(pt.getPlaintext());
    return w.finish();
}

CKKSEncoding ckks_deserialize_encoding(const CKKSContext& c, std::span<const uint8_t> bytes) {
    const auto& P = c.getParams();
    Reader r(bytes, kMagicPt, fingerprint(P), "ckks_deserialize_encoding");
    if (r.flags() > 1) r.fail("invalid flags");
    const uint64_t N = r.u64(), level = r.u64();
    if (N != P.getN()) r.fail("ring dimension mismatch");
    if (level > P.getMaxLevel()) r.fail("level exceeds the modulus chain");
    const long double scale = r.scale();
    auto w = r.words((level + 1) * N);
    check_residues(w, P.getModuli(), N, level + 1, r);
    r.done();
    return CKKSEncoding(c, std::move(w), scale, level, r.flags() == 1);
}

std::vector<uint8_t> ckks_serialize_evaluation_keys(const CKKSContext& c) {
    const auto& P = c.getParams();
    Writer w(kMagicKeys, (P.fold_hybrid_keyswitch ? 1u : 0u) | (P.montgomery ? 2u : 0u), fingerprint(P));
    const auto rots = c.getRotationKeyValues();
    const bool sparse = c.hasSparseSecretEncapsulationKeys();
    w.u64(c.keyDigitCount());
    w.u64(1 + (c.hasConjugationKey() ? 1 : 0) + rots.size() + (sparse ? 2 : 0));
    auto entry = [&](uint32_t kind, int64_t rot, const MyVector<uint64_t>& key) {
        w.u32(kind);
        w.u32(0);
        w.i64(rot);
        w.u64(key.size());
        w.words(key);
    };
    entry(kRelin, 0, c.downloadKey(c.getRelinKeyDevice()));
    if (c.hasConjugationKey()) entry(kConj, 0, c.downloadKey(c.getConjugationKeyDevice()));
    for (int64_t r : rots) entry(kRotation, r, c.downloadKey(c.getRotationKeyDevice(r)));
    if (sparse) {
        entry(kDenseToSparse, 0, c.downloadDenseToSparseSwitchKey());
        entry(kSparseToDense, 0, c.downloadKey(c.getSparseToDenseSwitchKeyDevice()));
    }
    return w.finish();
}

void ckks_load_evaluation_keys(CKKSContext& c, std::span<const uint8_t> bytes) {
    const auto& P = c.getParams();
    Reader r(bytes, kMagicKeys, fingerprint(P), "ckks_load_evaluation_keys");
    const uint32_t want_flags = (P.fold_hybrid_keyswitch ? 1u : 0u) | (P.montgomery ? 2u : 0u);
    if (r.flags() != want_flags) r.fail("key form (folding / Montgomery) differs from the context's");
    if (r.u64() != c.keyDigitCount()) r.fail("digit count differs from the context's");
    const uint64_t count = r.u64();
    const size_t N = P.getN();
    // parse everything before installing anything
    struct Entry {
        uint32_t kind;
        int64_t rot;
        MyVector<uint64_t> key;
    };
    std::vector<Entry> entries;
    for (uint64_t e = 0; e < count; ++e) {
        Entry x;
        x.kind = r.u32();
        if (r.u32() != 0) r.fail("invalid entry header");
        x.rot = r.i64();
        const uint64_t words = r.u64();
        const uint64_t expect = x.kind == kDenseToSparse ? 4 * N : c.switchKeyWords();
        if (x.kind > kSparseToDense) r.fail("unknown key kind");
        if (words != expect) r.fail("key size differs from the context's");
        x.key = r.words(words);
        entries.push_back(std::move(x));
    }
    r.done();
    for (auto& x : entries) {
        switch (x.kind) {
            case kRelin: c.setRelinKey(x.key); break;
            case kConj: c.setConjugationKey(x.key); break;
            case kRotation: c.setRotationKey(x.rot, x.key); break;
            case kDenseToSparse: c.setDenseToSparseSwitchKey(x.key); break;
            case kSparseToDense: c.setSparseToDenseSwitchKey(x.key); break;
        }
    }
}
