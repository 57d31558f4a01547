// Mitsuba ".serialized" mesh reader behind the C-ABI (dr_load_serialized,
// reference load_serialized.cpp:1-277, a pybind11 module there).
//
// File layout (Mitsuba 0.5 serialized format, versions 3 and 4): a sequence
// of shapes, each a 2-byte magic number and a 2-byte version followed by a
// zlib stream holding
//   uint32 flags (0x1 normals, 0x2 texture coordinates, 0x8 colors,
//                 0x2000 double precision),
//   version 4: the shape's name, NUL-terminated,
//   uint64 vertex count, uint64 triangle count,
//   positions [n][3], normals [n][3] if flagged, uvs [n][2] if flagged,
//   colors [n][3] if flagged (float or double), indices [m][3] (int32);
// the file ends with the shapes' byte offsets (version 4: uint64 each,
// version 3: uint32 each) and a uint32 shape count.  As the reference, shape
// 0 is read from just after the file's header, shape i > 0 from its offset
// in that table, with the first shape's version.
#include "../../include/delta_ray.h"

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

void dr_set_error(const std::string &msg);  // render.hip

namespace {

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

// zlib inflation of the stream that starts at the file's current position
class Inflater {
  public:
    explicit Inflater(FILE *f) : f_(f) {
        std::memset(&z_, 0, sizeof(z_));
        ok_ = inflateInit2(&z_, 15) == Z_OK;
    }
    ~Inflater() {
        if (ok_) inflateEnd(&z_);
    }
    bool ok() const { return ok_; }
    // reads exactly n bytes; false (with the reason in err) otherwise
    bool read(void *dst, size_t n, std::string &err) {
        z_.next_out = (Bytef *)dst;
        while (n > 0) {
            if (z_.avail_in == 0) {
                size_t got = std::fread(buf_, 1, sizeof(buf_), f_);
                if (got == 0) {
                    err = "read less data than expected";
                    return false;
                }
                z_.next_in = buf_;
                z_.avail_in = (uInt)got;
            }
            uInt chunk = (uInt)std::min<size_t>(n, 1u << 30);
            z_.avail_out = chunk;
            int r = inflate(&z_, Z_NO_FLUSH);
            if (r == Z_STREAM_ERROR || r == Z_NEED_DICT || r == Z_DATA_ERROR || r == Z_MEM_ERROR) {
                err = std::string("inflate(): ") + (z_.msg ? z_.msg : "error");
                return false;
            }
            size_t out = chunk - z_.avail_out;
            n -= out;
            if (n > 0 && r == Z_STREAM_END) {
                err = "inflate(): attempting to read past the end of the stream";
                return false;
            }
        }
        return true;
    }

  private:
    FILE *f_;
    z_stream z_;
    bool ok_ = false;
    unsigned char buf_[32768];
};

template <class P>
bool read_array(Inflater &zs, float *out, size_t count, std::string &err) {
    std::vector<P> tmp(std::min<size_t>(count, 1 << 20));
    size_t done = 0;
    while (done < count) {
        size_t k = std::min(tmp.size(), count - done);
        if (!zs.read(tmp.data(), k * sizeof(P), err)) return false;
        for (size_t i = 0; i < k; i++) out[done + i] = (float)tmp[i];
        done += k;
    }
    return true;
}

dr_status fail(const std::string &m) {
    dr_set_error(m);
    return DR_ERROR_INVALID_ARGUMENT;
}

}  // namespace

extern "C" {

void dr_mesh_data_free(dr_mesh_data *m) {
    if (!m) return;
    std::free(m->vertices);
    std::free(m->normals);
    std::free(m->uvs);
    std::free(m->colors);
    std::free(m->indices);
    std::memset(m, 0, sizeof(*m));
}

dr_status dr_load_serialized(const char *filename, int shape_index, dr_mesh_data *out) {
    if (!filename || !out || shape_index < 0) return fail("dr_load_serialized: bad arguments");
    std::memset(out, 0, sizeof(*out));
    File fh;
    fh.f = std::fopen(filename, "rb");
    if (!fh.f) return fail(std::string("cannot open ") + filename);
    FILE *f = fh.f;
    uint16_t magic = 0, version = 0;
    if (std::fread(&magic, 2, 1, f) != 1 || std::fread(&version, 2, 1, f) != 1) return fail("truncated header");
    if (version != 3 && version != 4) return fail("unsupported serialized version " + std::to_string(version));
    if (shape_index > 0) {
        uint32_t count = 0;
        if (std::fseek(f, -4, SEEK_END) != 0 || std::fread(&count, 4, 1, f) != 1) return fail("truncated shape table");
        if ((uint32_t)shape_index >= count)
            return fail("shape index " + std::to_string(shape_index) + " out of range (" + std::to_string(count) +
                        " shapes)");
        uint64_t offset = 0;
        if (version == 4) {
            if (std::fseek(f, -(long)(8 * (uint64_t)(count - shape_index)) - 4, SEEK_END) != 0 ||
                std::fread(&offset, 8, 1, f) != 1)
                return fail("truncated shape table");
        } else {
            uint32_t o = 0;
            if (std::fseek(f, -(long)(4 * (uint64_t)(count - shape_index + 1)), SEEK_END) != 0 ||
                std::fread(&o, 4, 1, f) != 1)
                return fail("truncated shape table");
            offset = o;
        }
        if (std::fseek(f, (long)offset + 4, SEEK_SET) != 0) return fail("bad shape offset");
    }
    Inflater zs(f);
    if (!zs.ok()) return fail("could not initialise zlib");
    std::string err;
    uint32_t flags = 0;
    if (!zs.read(&flags, 4, err)) return fail(err);
    if (version == 4) {
        char c = 1;
        while (c != '\0')
            if (!zs.read(&c, 1, err)) return fail(err);
    }
    uint64_t nv = 0, nt = 0;
    if (!zs.read(&nv, 8, err) || !zs.read(&nt, 8, err)) return fail(err);
    if (nv > (uint64_t)INT32_MAX / 3 || nt > (uint64_t)INT32_MAX / 3) return fail("mesh too large");
    const bool dbl = (flags & 0x2000) != 0;
    auto floats = [&](float **dst, size_t n) -> bool {
        *dst = (float *)std::malloc(std::max<size_t>(n, 1) * sizeof(float));
        if (!*dst) {
            err = "host allocation failed";
            return false;
        }
        return dbl ? read_array<double>(zs, *dst, n, err) : read_array<float>(zs, *dst, n, err);
    };
    out->num_vertices = (int)nv;
    out->num_triangles = (int)nt;
    bool ok = floats(&out->vertices, 3 * nv);
    if (ok && (flags & 0x1)) ok = floats(&out->normals, 3 * nv);
    if (ok && (flags & 0x2)) ok = floats(&out->uvs, 2 * nv);
    if (ok && (flags & 0x8)) ok = floats(&out->colors, 3 * nv);
    if (ok) {
        out->indices = (int *)std::malloc(std::max<size_t>(3 * nt, 1) * sizeof(int));
        ok = out->indices != nullptr && zs.read(out->indices, 3 * nt * sizeof(int), err);
        if (!out->indices) err = "host allocation failed";
    }
    if (!ok) {
        dr_mesh_data_free(out);
        return fail(err);
    }
    return DR_OK;
}

}  // extern "C"
