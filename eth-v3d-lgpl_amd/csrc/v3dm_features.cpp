// SIFT feature archives (".features" files) read and written on the host.
//
// Format of V3D::serializeDataToFile(name, SerializableVector<SIFT_Feature>)
// (V3D/Base/v3d_serialization.h:953-1017, 851-859, 1042-1059; record V3D/Base/v3d_feature.h:
// 41-49):
//   binary: "V3DBIN", uint32 count, then per feature int32 id, float position x, y,
//           direction x, y, scale, magnitude, 128 float descriptor -- 540 bytes, integers
//           little-endian byte by byte (v3d_serialization.h:415-432), floats as their
//           IEEE bytes (little-endian on every platform the library runs on)
//   text:   "V3DTXT", then whitespace-separated tokens: count "{" and per feature
//           "SIFT_Feature" "{" id x y dx dy scale magnitude d0 .. d127 "}", then "}"
//           (TextOStreamArchive, v3d_serialization.h:176-260: values printed with the
//           stream's default 6 significant digits)
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/v3dm.h"
#include "v3dm_ext.h"

namespace {

constexpr int kTag = 6;
constexpr int kRecordFloats = 6 + 128;  // position, direction, scale, magnitude, descriptor
constexpr size_t kRecordBytes = 4 + kRecordFloats * 4;

int ioFail(const std::string& msg) { return v3dm::ctx_fail(nullptr, V3DM_ERR_IO, msg.c_str()); }

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) fclose(f);
    }
};

uint32_t le32(unsigned char const* b)
{
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
}

void put32(unsigned char* b, uint32_t v)
{
    b[0] = (unsigned char)(v & 0xff);
    b[1] = (unsigned char)((v >> 8) & 0xff);
    b[2] = (unsigned char)((v >> 16) & 0xff);
    b[3] = (unsigned char)((v >> 24) & 0xff);
}

float lef(unsigned char const* b)
{
    uint32_t const u = le32(b);
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// Whitespace tokenizer over a text archive.
struct Tokens {
    FILE* f;
    std::string tok;
    bool next()
    {
        tok.clear();
        int c;
        do {
            c = fgetc(f);
        } while (c != EOF && (c == ' ' || c == '\n' || c == '\t' || c == '\r'));
        while (c != EOF && !(c == ' ' || c == '\n' || c == '\t' || c == '\r')) {
            tok.push_back((char)c);
            c = fgetc(f);
        }
        return !tok.empty();
    }
    bool expect(const char* s) { return next() && tok == s; }
    bool number(double& v)
    {
        if (!next()) return false;
        char* end = nullptr;
        errno = 0;
        v = strtod(tok.c_str(), &end);
        return end && *end == 0;
    }
};

struct Header {
    bool binary = true;
    uint32_t count = 0;
};

int openArchive(const char* path, File& fh, Header& hd, Tokens& tk)
{
    if (!path) return ioFail("null path");
    fh.f = fopen(path, "rb");
    if (!fh.f) return ioFail(std::string("cannot open ") + path + ": " + strerror(errno));
    char tag[kTag];
    if (fread(tag, 1, kTag, fh.f) != (size_t)kTag) return ioFail(std::string(path) + ": shorter than the archive tag");
    if (memcmp(tag, "V3DBIN", kTag) == 0) {
        hd.binary = true;
        unsigned char b[4];
        if (fread(b, 1, 4, fh.f) != 4) return ioFail(std::string(path) + ": missing feature count");
        hd.count = le32(b);
    } else if (memcmp(tag, "V3DTXT", kTag) == 0) {
        hd.binary = false;
        tk.f = fh.f;
        double v;
        if (!tk.number(v) || v < 0 || v > 4294967295.0 || v != std::floor(v))
            return ioFail(std::string(path) + ": bad feature count");
        hd.count = (uint32_t)v;
        if (!tk.expect("{")) return ioFail(std::string(path) + ": expected '{'");
    } else {
        return ioFail(std::string(path) + ": unknown archive magic tag");  // v3d_serialization.h:1000
    }
    return V3DM_OK;
}

}  // namespace

extern "C" {

int v3dm_features_count(const char* path, int* nFeatures, int* binary)
{
    File fh;
    Header hd;
    Tokens tk{nullptr, {}};
    int const rc = openArchive(path, fh, hd, tk);
    if (rc) return rc;
    if (hd.count > 0x7fffffffu) return ioFail(std::string(path) + ": feature count above 2^31 - 1");
    if (nFeatures) *nFeatures = (int)hd.count;
    if (binary) *binary = hd.binary ? 1 : 0;
    return V3DM_OK;
}

int v3dm_features_read(const char* path, int cap, int32_t* ids, float* xy, float* dir, float* scale,
                       float* magnitude, float* desc, int* n)
{
    if (cap < 0) return v3dm::ctx_fail(nullptr, V3DM_ERR_INVALID, "negative capacity");
    File fh;
    Header hd;
    Tokens tk{nullptr, {}};
    int rc = openArchive(path, fh, hd, tk);
    if (rc) return rc;
    int const count = (int)std::min<uint32_t>(hd.count, 0x7fffffffu);
    int const take = std::min(count, cap);
    float rec[kRecordFloats];
    std::vector<unsigned char> buf(kRecordBytes);
    for (int k = 0; k < take; ++k) {
        int32_t id = 0;
        if (hd.binary) {
            if (fread(buf.data(), 1, kRecordBytes, fh.f) != kRecordBytes)
                return ioFail(std::string(path) + ": truncated at feature " + std::to_string(k));
            id = (int32_t)le32(buf.data());
            for (int t = 0; t < kRecordFloats; ++t) rec[t] = lef(buf.data() + 4 + 4 * t);
        } else {
            double v;
            if (!tk.expect("SIFT_Feature") || !tk.expect("{"))
                return ioFail(std::string(path) + ": expected 'SIFT_Feature {' at feature " + std::to_string(k));
            if (!tk.number(v)) return ioFail(std::string(path) + ": bad id at feature " + std::to_string(k));
            id = (int32_t)v;
            for (int t = 0; t < kRecordFloats; ++t) {
                if (!tk.number(v))
                    return ioFail(std::string(path) + ": bad value at feature " + std::to_string(k));
                rec[t] = (float)v;
            }
            if (!tk.expect("}")) return ioFail(std::string(path) + ": expected '}' at feature " + std::to_string(k));
        }
        if (ids) ids[k] = id;
        if (xy) {
            xy[2 * k] = rec[0];
            xy[2 * k + 1] = rec[1];
        }
        if (dir) {
            dir[2 * k] = rec[2];
            dir[2 * k + 1] = rec[3];
        }
        if (scale) scale[k] = rec[4];
        if (magnitude) magnitude[k] = rec[5];
        if (desc) memcpy(desc + (size_t)k * 128, rec + 6, 128 * sizeof(float));
    }
    if (n) *n = take;
    return V3DM_OK;
}

int v3dm_features_write(const char* path, int n, const int32_t* ids, const float* xy, const float* dir,
                        const float* scale, const float* magnitude, const float* desc, int binary)
{
    if (!path) return ioFail("null path");
    if (n < 0) return v3dm::ctx_fail(nullptr, V3DM_ERR_INVALID, "negative feature count");
    if (n > 0 && !desc) return v3dm::ctx_fail(nullptr, V3DM_ERR_INVALID, "null descriptor array");
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) return ioFail(std::string("cannot create ") + path + ": " + strerror(errno));
    auto rec = [&](int k, float* r) {
        r[0] = xy ? xy[2 * k] : 0.0f;
        r[1] = xy ? xy[2 * k + 1] : 0.0f;
        r[2] = dir ? dir[2 * k] : 0.0f;
        r[3] = dir ? dir[2 * k + 1] : 0.0f;
        r[4] = scale ? scale[k] : 0.0f;
        r[5] = magnitude ? magnitude[k] : 0.0f;
        memcpy(r + 6, desc + (size_t)k * 128, 128 * sizeof(float));
    };
    float r[kRecordFloats];
    bool ok = true;
    if (binary) {
        unsigned char head[kTag + 4];
        memcpy(head, "V3DBIN", kTag);
        put32(head + kTag, (uint32_t)n);
        ok = fwrite(head, 1, sizeof(head), fh.f) == sizeof(head);
        std::vector<unsigned char> buf(kRecordBytes);
        for (int k = 0; ok && k < n; ++k) {
            put32(buf.data(), (uint32_t)(ids ? ids[k] : k));
            rec(k, r);
            for (int t = 0; t < kRecordFloats; ++t) {
                uint32_t u;
                memcpy(&u, &r[t], 4);
                put32(buf.data() + 4 + 4 * t, u);
            }
            ok = fwrite(buf.data(), 1, kRecordBytes, fh.f) == kRecordBytes;
        }
    } else {
        // the text archive's layout: scopes indent by two, one feature per line
        ok = fprintf(fh.f, "V3DTXT\n%d { ", n) > 0;
        for (int k = 0; ok && k < n; ++k) {
            rec(k, r);
            ok = fprintf(fh.f, "SIFT_Feature { %d ", ids ? ids[k] : k) > 0;
            for (int t = 0; ok && t < kRecordFloats; ++t) ok = fprintf(fh.f, "%g ", (double)r[t]) > 0;
            ok = ok && fprintf(fh.f, "} \n  ") > 0;
        }
        ok = ok && fprintf(fh.f, "} \n") > 0;
    }
    if (!ok) return ioFail(std::string("write failed: ") + path);
    if (fflush(fh.f) != 0) return ioFail(std::string("write failed: ") + path);
    return V3DM_OK;
}

}  // extern "C"
