// On-disk page input: a TIFF page reader that decodes grey pages straight into caller memory
// (pinned host buffers, so the upload to HBM can overlap the next page's read and the
// previous page's detection). Host code only.
//
// The reference reads pages with tifffile (subpx/mosaic.py:51-79: every page of every file of
// a stripe/board, in file order, as page.asarray()). This reader covers what such stripe
// pages are: one grey sample per pixel, 8- or 16-bit unsigned, classic TIFF or BigTIFF in
// either byte order, strips or tiles, uncompressed, PackBits or Deflate (zlib), horizontal
// differencing predictor or none. The values are the stored samples (no photometric
// inversion), as tifffile's asarray returns them. Anything else is refused with
// SPX_E_UNSUPPORTED and a message naming the tag.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spx_internal.hpp"

namespace {

struct PageDesc {
    int64_t H = 0, W = 0;
    int bits = 0;
    int compression = 1;
    int predictor = 1;
    bool tiled = false;
    int64_t rows_per_strip = 0;      // strips
    int64_t tile_w = 0, tile_h = 0;  // tiles
    std::vector<uint64_t> offsets, counts;
};

}  // namespace

struct SpxTiff {
    FILE* fp = nullptr;
    std::string path;
    bool big_endian = false;
    bool bigtiff = false;
    std::vector<PageDesc> pages;
    std::vector<unsigned char> chunk;  // one strip / tile, compressed
    std::vector<unsigned char> plain;  // the same, decoded
};

namespace {

constexpr int kMaxPages = 1 << 20;

struct Reader {
    SpxTiff* t;
    bool ok = true;
    std::string err;

    bool read_at(uint64_t off, void* dst, size_t n) {
        if (fseeko(t->fp, (off_t)off, SEEK_SET) != 0 || fread(dst, 1, n, t->fp) != n) {
            fail("read past the end of the file at offset " + std::to_string(off));
            return false;
        }
        return true;
    }
    void fail(const std::string& m) {
        if (ok) err = m;
        ok = false;
    }
    uint64_t get(const unsigned char* p, int n) const {
        uint64_t v = 0;
        if (t->big_endian)
            for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
        else
            for (int i = n - 1; i >= 0; --i) v = (v << 8) | p[i];
        return v;
    }
};

int type_size(int type) {
    switch (type) {
        case 1: case 2: case 6: case 7: return 1;  // BYTE ASCII SBYTE UNDEFINED
        case 3: case 8: return 2;                  // SHORT SSHORT
        case 4: case 9: case 11: case 13: return 4;  // LONG SLONG FLOAT IFD
        case 5: case 10: case 12: case 16: case 17: case 18: return 8;  // RATIONAL DOUBLE LONG8 SLONG8 IFD8
        default: return 0;
    }
}

// Values of an integer-typed entry (BYTE, SHORT, LONG, LONG8, IFD, IFD8).
bool entry_values(Reader& r, const unsigned char* e, int entry_bytes, std::vector<uint64_t>& out) {
    const int type = (int)r.get(e + 2, 2);
    const int sz = type_size(type);
    const bool intlike = type == 1 || type == 3 || type == 4 || type == 16 || type == 13 || type == 18;
    if (!intlike) {
        r.fail("TIFF tag " + std::to_string(r.get(e, 2)) + " has non-integer type " + std::to_string(type));
        return false;
    }
    const int cnt_bytes = r.t->bigtiff ? 8 : 4;
    const uint64_t count = r.get(e + 4, cnt_bytes);
    const int inline_bytes = r.t->bigtiff ? 8 : 4;
    if (count > (uint64_t)1 << 28) {
        r.fail("TIFF tag with an implausible count");
        return false;
    }
    std::vector<unsigned char> raw((size_t)(count * sz));
    const unsigned char* val = e + 4 + cnt_bytes;
    if (count * sz <= (uint64_t)inline_bytes) {
        if (!raw.empty()) memcpy(raw.data(), val, raw.size());
    } else if (!r.read_at(r.get(val, inline_bytes), raw.data(), raw.size())) {
        return false;
    }
    out.resize((size_t)count);
    for (uint64_t i = 0; i < count; ++i) out[(size_t)i] = r.get(raw.data() + i * sz, sz);
    (void)entry_bytes;
    return true;
}

int parse(SpxTiff* t) {
    Reader r{t};
    unsigned char head[16];
    if (fread(head, 1, 8, t->fp) != 8) {
        spx::set_error("TIFF: " + t->path + " is shorter than a TIFF header");
        return SPX_E_UNSUPPORTED;
    }
    if (head[0] == 'I' && head[1] == 'I') t->big_endian = false;
    else if (head[0] == 'M' && head[1] == 'M') t->big_endian = true;
    else {
        spx::set_error("TIFF: " + t->path + " has no TIFF byte-order mark");
        return SPX_E_UNSUPPORTED;
    }
    const int magic = (int)r.get(head + 2, 2);
    uint64_t ifd = 0;
    if (magic == 42) {
        t->bigtiff = false;
        ifd = r.get(head + 4, 4);
    } else if (magic == 43) {
        t->bigtiff = true;
        if (fread(head + 8, 1, 8, t->fp) != 8 || r.get(head + 4, 2) != 8) {
            spx::set_error("TIFF: " + t->path + " has a malformed BigTIFF header");
            return SPX_E_UNSUPPORTED;
        }
        ifd = r.get(head + 8, 8);
    } else {
        spx::set_error("TIFF: " + t->path + " has version " + std::to_string(magic) + " (not 42 or 43)");
        return SPX_E_UNSUPPORTED;
    }
    const int entry_bytes = t->bigtiff ? 20 : 12;
    const int count_bytes = t->bigtiff ? 8 : 2;
    const int next_bytes = t->bigtiff ? 8 : 4;
    std::vector<uint64_t> seen;
    while (ifd != 0 && r.ok) {
        for (uint64_t s : seen)
            if (s == ifd) {
                spx::set_error("TIFF: " + t->path + " has a loop in its page chain");
                return SPX_E_UNSUPPORTED;
            }
        seen.push_back(ifd);
        if ((int)seen.size() > kMaxPages) {
            spx::set_error("TIFF: " + t->path + " has more pages than the reader accepts");
            return SPX_E_UNSUPPORTED;
        }
        unsigned char cb[8];
        if (!r.read_at(ifd, cb, count_bytes)) break;
        const uint64_t n = r.get(cb, count_bytes);
        if (n > 4096) {
            spx::set_error("TIFF: " + t->path + " has an IFD of " + std::to_string(n) + " entries");
            return SPX_E_UNSUPPORTED;
        }
        std::vector<unsigned char> ents((size_t)(n * entry_bytes + next_bytes));
        if (!r.read_at(ifd + count_bytes, ents.data(), ents.size())) break;
        PageDesc pd;
        int spp = 1, planar = 1, sample_format = 1;
        for (uint64_t i = 0; i < n && r.ok; ++i) {
            const unsigned char* e = ents.data() + i * entry_bytes;
            const int tag = (int)r.get(e, 2);
            std::vector<uint64_t> v;
            switch (tag) {
                case 256: case 257: case 258: case 259: case 273: case 277: case 278: case 279: case 284:
                case 317: case 322: case 323: case 324: case 325: case 339:
                    if (!entry_values(r, e, entry_bytes, v) || v.empty()) {
                        r.fail("TIFF tag " + std::to_string(tag) + " has no values");
                        break;
                    }
                    if (tag == 256) pd.W = (int64_t)v[0];
                    else if (tag == 257) pd.H = (int64_t)v[0];
                    else if (tag == 258) {
                        pd.bits = (int)v[0];
                        for (uint64_t b : v)
                            if ((int)b != pd.bits) r.fail("TIFF: samples of different bit depths");
                    } else if (tag == 259) pd.compression = (int)v[0];
                    else if (tag == 273 || tag == 324) pd.offsets = v;
                    else if (tag == 279 || tag == 325) pd.counts = v;
                    else if (tag == 277) spp = (int)v[0];
                    else if (tag == 278) pd.rows_per_strip = (int64_t)v[0];
                    else if (tag == 284) planar = (int)v[0];
                    else if (tag == 317) pd.predictor = (int)v[0];
                    else if (tag == 322) pd.tile_w = (int64_t)v[0];
                    else if (tag == 323) pd.tile_h = (int64_t)v[0];
                    else if (tag == 339) sample_format = (int)v[0];
                    if (tag == 322 || tag == 323 || tag == 324 || tag == 325) pd.tiled = true;
                    break;
                default:
                    break;
            }
        }
        if (!r.ok) break;
        const std::string where = t->path + " page " + std::to_string(t->pages.size());
        if (pd.H <= 0 || pd.W <= 0) {
            spx::set_error("TIFF: " + where + " has no image size");
            return SPX_E_UNSUPPORTED;
        }
        if (spp != 1 || (pd.bits != 8 && pd.bits != 16) || sample_format != 1) {
            spx::set_error("TIFF: " + where + " is not a grey page of 8- or 16-bit unsigned samples (samples " +
                           std::to_string(spp) + ", bits " + std::to_string(pd.bits) + ", format " +
                           std::to_string(sample_format) + ")");
            return SPX_E_UNSUPPORTED;
        }
        (void)planar;  // one sample per pixel: both planar configurations store the same bytes
        if (pd.compression != 1 && pd.compression != 32773 && pd.compression != 8 && pd.compression != 32946) {
            spx::set_error("TIFF: " + where + " uses compression " + std::to_string(pd.compression) +
                           " (none, PackBits and Deflate are read)");
            return SPX_E_UNSUPPORTED;
        }
        if (pd.predictor != 1 && pd.predictor != 2) {
            spx::set_error("TIFF: " + where + " uses predictor " + std::to_string(pd.predictor));
            return SPX_E_UNSUPPORTED;
        }
        int64_t chunks = 0;
        if (pd.tiled) {
            if (pd.tile_w <= 0 || pd.tile_h <= 0) {
                spx::set_error("TIFF: " + where + " has tiles without a tile size");
                return SPX_E_UNSUPPORTED;
            }
            chunks = ((pd.W + pd.tile_w - 1) / pd.tile_w) * ((pd.H + pd.tile_h - 1) / pd.tile_h);
        } else {
            if (pd.rows_per_strip <= 0 || pd.rows_per_strip > pd.H) pd.rows_per_strip = pd.H;
            chunks = (pd.H + pd.rows_per_strip - 1) / pd.rows_per_strip;
        }
        if ((int64_t)pd.offsets.size() < chunks || (int64_t)pd.counts.size() < chunks) {
            spx::set_error("TIFF: " + where + " lists fewer strips or tiles than its size needs");
            return SPX_E_UNSUPPORTED;
        }
        t->pages.push_back(pd);
        ifd = r.get(ents.data() + n * entry_bytes, next_bytes);
    }
    if (!r.ok) {
        spx::set_error("TIFF: " + t->path + ": " + r.err);
        return SPX_E_UNSUPPORTED;
    }
    return SPX_OK;
}

// PackBits: a header byte n; 0..127 copies n + 1 bytes, -1..-127 repeats the next byte 1 - n
// times, -128 is skipped.
bool unpack_bits(const unsigned char* src, size_t n, unsigned char* dst, size_t want) {
    size_t i = 0, o = 0;
    while (i < n && o < want) {
        const int h = (int)(signed char)src[i++];
        if (h >= 0) {
            const size_t c = (size_t)h + 1;
            if (i + c > n || o + c > want) return false;
            memcpy(dst + o, src + i, c);
            i += c;
            o += c;
        } else if (h != -128) {
            const size_t c = (size_t)(1 - h);
            if (i >= n || o + c > want) return false;
            memset(dst + o, src[i++], c);
            o += c;
        }
    }
    return o == want;
}

// One strip or tile of `rows` x `cols` samples decoded into t->plain (native byte order,
// predictor undone).
int decode_chunk(SpxTiff* t, const PageDesc& pd, size_t k, int64_t rows, int64_t cols) {
    const int bps = pd.bits / 8;
    const size_t want = (size_t)(rows * cols * bps);
    t->chunk.resize((size_t)pd.counts[k]);
    if (fseeko(t->fp, (off_t)pd.offsets[k], SEEK_SET) != 0 ||
        fread(t->chunk.data(), 1, t->chunk.size(), t->fp) != t->chunk.size()) {
        spx::set_error("TIFF: " + t->path + ": a strip or tile lies past the end of the file");
        return SPX_E_UNSUPPORTED;
    }
    t->plain.resize(want);
    if (pd.compression == 1) {
        if (t->chunk.size() < want) {
            spx::set_error("TIFF: " + t->path + ": an uncompressed strip or tile is short");
            return SPX_E_UNSUPPORTED;
        }
        memcpy(t->plain.data(), t->chunk.data(), want);
    } else if (pd.compression == 32773) {
        if (!unpack_bits(t->chunk.data(), t->chunk.size(), t->plain.data(), want)) {
            spx::set_error("TIFF: " + t->path + ": a PackBits strip or tile does not decode to its size");
            return SPX_E_UNSUPPORTED;
        }
    } else {
        uLongf got = (uLongf)want;
        const int zr = uncompress(t->plain.data(), &got, t->chunk.data(), (uLong)t->chunk.size());
        if (zr != Z_OK || got != (uLongf)want) {
            spx::set_error("TIFF: " + t->path + ": a Deflate strip or tile does not decode to its size");
            return SPX_E_UNSUPPORTED;
        }
    }
    if (bps == 2 && t->big_endian)
        for (size_t i = 0; i + 1 < want; i += 2) std::swap(t->plain[i], t->plain[i + 1]);
    if (pd.predictor == 2) {
        for (int64_t y = 0; y < rows; ++y) {
            if (bps == 1) {
                unsigned char* p = t->plain.data() + y * cols;
                for (int64_t x = 1; x < cols; ++x) p[x] = (unsigned char)(p[x] + p[x - 1]);
            } else {
                uint16_t* p = reinterpret_cast<uint16_t*>(t->plain.data()) + y * cols;
                for (int64_t x = 1; x < cols; ++x) p[x] = (uint16_t)(p[x] + p[x - 1]);
            }
        }
    }
    return SPX_OK;
}

}  // namespace

extern "C" {

int spx_tiff_open(const char* path, spx_tiff_t* out) {
    spx::clear_error();
    SPX_REQUIRE(path != nullptr && out != nullptr, "spx_tiff_open: null pointer");
    *out = nullptr;
    SpxTiff* t = new SpxTiff();
    t->path = path;
    t->fp = fopen(path, "rb");
    if (t->fp == nullptr) {
        delete t;
        spx::set_error(std::string("spx_tiff_open: cannot open ") + path);
        return SPX_E_ARG;
    }
    const int rc = parse(t);
    if (rc != SPX_OK) {
        fclose(t->fp);
        delete t;
        return rc;
    }
    *out = t;
    return SPX_OK;
}

int spx_tiff_page_count(spx_tiff_t t, int32_t* n_out) {
    spx::clear_error();
    SPX_REQUIRE(t != nullptr && n_out != nullptr, "spx_tiff_page_count: null pointer");
    *n_out = (int32_t)t->pages.size();
    return SPX_OK;
}

int spx_tiff_page_info(spx_tiff_t t, int32_t page, int64_t* H, int64_t* W, int32_t* bits) {
    spx::clear_error();
    SPX_REQUIRE(t != nullptr && H != nullptr && W != nullptr && bits != nullptr, "spx_tiff_page_info: null pointer");
    SPX_REQUIRE(page >= 0 && page < (int32_t)t->pages.size(), "spx_tiff_page_info: page out of range");
    const PageDesc& pd = t->pages[(size_t)page];
    *H = pd.H;
    *W = pd.W;
    *bits = pd.bits;
    return SPX_OK;
}

int spx_tiff_read_page(spx_tiff_t t, int32_t page, void* dst, int64_t dst_pitch_bytes) {
    spx::clear_error();
    SPX_REQUIRE(t != nullptr && dst != nullptr, "spx_tiff_read_page: null pointer");
    SPX_REQUIRE(page >= 0 && page < (int32_t)t->pages.size(), "spx_tiff_read_page: page out of range");
    const PageDesc& pd = t->pages[(size_t)page];
    const int bps = pd.bits / 8;
    SPX_REQUIRE(dst_pitch_bytes >= pd.W * bps, "spx_tiff_read_page: row pitch below the page's row bytes");
    unsigned char* out = static_cast<unsigned char*>(dst);
    if (!pd.tiled) {
        const int64_t nstrips = (pd.H + pd.rows_per_strip - 1) / pd.rows_per_strip;
        for (int64_t s = 0; s < nstrips; ++s) {
            const int64_t y0 = s * pd.rows_per_strip;
            const int64_t rows = (y0 + pd.rows_per_strip <= pd.H) ? pd.rows_per_strip : pd.H - y0;
            const int rc = decode_chunk(t, pd, (size_t)s, rows, pd.W);
            if (rc != SPX_OK) return rc;
            for (int64_t y = 0; y < rows; ++y)
                memcpy(out + (y0 + y) * dst_pitch_bytes, t->plain.data() + y * pd.W * bps, (size_t)(pd.W * bps));
        }
        return SPX_OK;
    }
    const int64_t across = (pd.W + pd.tile_w - 1) / pd.tile_w;
    const int64_t down = (pd.H + pd.tile_h - 1) / pd.tile_h;
    for (int64_t ty = 0; ty < down; ++ty)
        for (int64_t tx = 0; tx < across; ++tx) {
            // a tile is stored whole (tile_h x tile_w) even where it overhangs the page
            const int rc = decode_chunk(t, pd, (size_t)(ty * across + tx), pd.tile_h, pd.tile_w);
            if (rc != SPX_OK) return rc;
            const int64_t x0 = tx * pd.tile_w, y0 = ty * pd.tile_h;
            const int64_t w = (x0 + pd.tile_w <= pd.W) ? pd.tile_w : pd.W - x0;
            const int64_t h = (y0 + pd.tile_h <= pd.H) ? pd.tile_h : pd.H - y0;
            for (int64_t y = 0; y < h; ++y)
                memcpy(out + (y0 + y) * dst_pitch_bytes + x0 * bps, t->plain.data() + y * pd.tile_w * bps,
                       (size_t)(w * bps));
        }
    return SPX_OK;
}

int spx_tiff_close(spx_tiff_t t) {
    spx::clear_error();
    if (t == nullptr) return SPX_OK;
    if (t->fp != nullptr) fclose(t->fp);
    delete t;
    return SPX_OK;
}

}  // extern "C"
