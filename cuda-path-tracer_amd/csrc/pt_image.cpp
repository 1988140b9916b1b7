// pt_image.cpp -- texture image decoding for OBJ/MTL import and the PPM writer.
//
// The reference loads textures with the vendored stb_image
// (src/textures/texture.h:65-72, stbi_load(..., STBI_rgb)) and writes its
// output as ASCII PPM (src/app/main.cpp:40-46, src/display/color.h:6-11).
// This file decodes the lossless formats bit for bit as that loader returns
// them, 8-bit RGB rows top to bottom:
//   PNG (RFC 2083): all colour types and bit depths, Adam7, with its own
//                   DEFLATE decoder (RFC 1951); 16-bit samples keep their high
//                   byte, sub-byte grey is scaled by 255/(2^depth-1), alpha is
//                   dropped;
//   binary PNM:     P5 (grey, replicated) and P6, maxval <= 255;
//   JPEG:           baseline and progressive Huffman, 8-bit (see below).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "pt_scene_impl.h"

namespace {

[[noreturn]] void corrupt(const char *what) { throw std::runtime_error(what); }

// ---------------------------------------------------------------- DEFLATE
class bit_reader {
public:
    bit_reader(const uint8_t *p, size_t n) : data(p), size(n) {}
    uint32_t bits(int count) {
        uint32_t value = 0;
        for(int i = 0; i < count; i++) value |= uint32_t(bit()) << i;
        return value;
    }
    int bit() {
        if(pos >= size) corrupt("truncated deflate stream");
        int b = (data[pos] >> shift) & 1;
        if(++shift == 8) { shift = 0; pos++; }
        return b;
    }
    void align() { if(shift) { shift = 0; pos++; } }
    uint8_t byte() { if(pos >= size) corrupt("truncated deflate stream"); return data[pos++]; }
private:
    const uint8_t *data;
    size_t size, pos = 0;
    int shift = 0;
};

// Canonical Huffman code: count[len] codes of each length, symbols in code order.
struct huffman {
    uint16_t count[16] = {};
    std::vector<uint16_t> symbol;
    void build(const uint8_t *lengths, int n) {
        std::fill(std::begin(count), std::end(count), 0);
        for(int i = 0; i < n; i++) count[lengths[i]]++;
        count[0] = 0;
        int left = 1;
        for(int len = 1; len < 16; len++) {
            left = 2 * left - count[len];
            if(left < 0) corrupt("over-subscribed Huffman code");
        }
        uint16_t offs[16] = {};
        for(int len = 1; len < 15; len++) offs[len + 1] = uint16_t(offs[len] + count[len]);
        symbol.assign(size_t(n), 0);
        for(int i = 0; i < n; i++) if(lengths[i]) symbol[offs[lengths[i]]++] = uint16_t(i);
    }
    int decode(bit_reader& in) const {
        int code = 0, first = 0, index = 0;
        for(int len = 1; len < 16; len++) {
            code |= in.bit();
            int c = count[len];
            if(code - c < first) return symbol[size_t(index + (code - first))];
            index += c; first += c;
            first <<= 1; code <<= 1;
        }
        corrupt("invalid Huffman code");
    }
};

const uint16_t length_base[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                                  35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t length_extra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t dist_base[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385,
                                513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t dist_extra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

void inflate_block(bit_reader& in, const huffman& lit, const huffman& dist, std::vector<uint8_t>& out) {
    for(;;) {
        int sym = lit.decode(in);
        if(sym < 256) { out.push_back(uint8_t(sym)); continue; }
        if(sym == 256) return;
        sym -= 257;
        if(sym >= 29) corrupt("invalid deflate length symbol");
        size_t len = length_base[sym] + in.bits(length_extra[sym]);
        int ds = dist.decode(in);
        if(ds >= 30) corrupt("invalid deflate distance symbol");
        size_t d = dist_base[ds] + in.bits(dist_extra[ds]);
        if(d > out.size()) corrupt("deflate distance too far back");
        size_t from = out.size() - d;
        for(size_t i = 0; i < len; i++) out.push_back(out[from + i]);
    }
}

std::vector<uint8_t> inflate(const uint8_t *data, size_t size, size_t expected) {
    bit_reader in(data, size);
    std::vector<uint8_t> out;
    // DEFLATE expands at most 1032-fold (258-byte copies in 2 bits): no more
    // is reserved than the stream can produce.
    out.reserve(std::min(expected, size * 1032 + 1024));
    huffman fixed_lit, fixed_dist;
    bool fixed_ready = false;
    int last = 0;
    while(!last) {
        last = int(in.bits(1));
        int type = int(in.bits(2));
        if(type == 0) {
            in.align();
            uint16_t len = uint16_t(in.byte() | (in.byte() << 8));
            uint16_t nlen = uint16_t(in.byte() | (in.byte() << 8));
            if(uint16_t(~nlen) != len) corrupt("corrupt stored deflate block");
            for(int i = 0; i < len; i++) out.push_back(in.byte());
        } else if(type == 1) {
            if(!fixed_ready) {
                uint8_t lengths[288];
                for(int i = 0; i < 144; i++) lengths[i] = 8;
                for(int i = 144; i < 256; i++) lengths[i] = 9;
                for(int i = 256; i < 280; i++) lengths[i] = 7;
                for(int i = 280; i < 288; i++) lengths[i] = 8;
                fixed_lit.build(lengths, 288);
                uint8_t dl[30];
                std::fill(std::begin(dl), std::end(dl), uint8_t(5));
                fixed_dist.build(dl, 30);
                fixed_ready = true;
            }
            inflate_block(in, fixed_lit, fixed_dist, out);
        } else if(type == 2) {
            int nlen = int(in.bits(5)) + 257, ndist = int(in.bits(5)) + 1, ncode = int(in.bits(4)) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t code_lengths[19] = {};
            for(int i = 0; i < ncode; i++) code_lengths[order[i]] = uint8_t(in.bits(3));
            huffman code;
            code.build(code_lengths, 19);
            uint8_t lengths[320] = {};
            int n = 0;
            while(n < nlen + ndist) {
                int sym = code.decode(in);
                if(sym < 16) { lengths[n++] = uint8_t(sym); continue; }
                int repeat = 0;
                uint8_t value = 0;
                if(sym == 16) {
                    if(n == 0) corrupt("deflate repeat with no previous length");
                    value = lengths[n - 1]; repeat = 3 + int(in.bits(2));
                } else if(sym == 17) repeat = 3 + int(in.bits(3));
                else repeat = 11 + int(in.bits(7));
                if(n + repeat > nlen + ndist) corrupt("deflate code lengths overflow");
                while(repeat--) lengths[n++] = value;
            }
            huffman lit, dist;
            lit.build(lengths, nlen);
            dist.build(lengths + nlen, ndist);
            inflate_block(in, lit, dist, out);
        } else {
            corrupt("invalid deflate block type");
        }
    }
    return out;
}

// ---------------------------------------------------------------- PNG
uint32_t be32(const uint8_t *p) { return uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3]; }

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if(pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Undo the per-row filters of one (sub)image in place; returns rows of `stride` bytes.
void unfilter(uint8_t *rows, size_t height, size_t stride, size_t bpp, std::vector<uint8_t>& out) {
    out.assign(height * stride, 0);
    for(size_t y = 0; y < height; y++) {
        const uint8_t *src = rows + y * (stride + 1);
        int filter = src[0];
        src++;
        uint8_t *dst = out.data() + y * stride;
        const uint8_t *up = y ? dst - stride : nullptr;
        for(size_t i = 0; i < stride; i++) {
            int a = i >= bpp ? dst[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int x = src[i];
            switch(filter) {
                case 0: break;
                case 1: x += a; break;
                case 2: x += b; break;
                case 3: x += (a + b) >> 1; break;
                case 4: x += paeth(a, b, c); break;
                default: corrupt("invalid PNG filter");
            }
            dst[i] = uint8_t(x);
        }
    }
}

std::vector<uint8_t> decode_png(const std::vector<uint8_t>& file, int& width, int& height) {
    static const uint8_t signature[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if(file.size() < 8 || std::memcmp(file.data(), signature, 8) != 0) corrupt("not a PNG file");
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, color = -1, interlace = 0;
    std::vector<uint8_t> palette, idat;
    bool header = false, ended = false;
    while(!ended) {
        if(pos + 12 > file.size()) corrupt("truncated PNG chunk");
        uint32_t len = be32(&file[pos]);
        const uint8_t *type = &file[pos + 4];
        if(size_t(len) > file.size() - pos - 12) corrupt("truncated PNG chunk");
        const uint8_t *body = &file[pos + 8];
        if(!std::memcmp(type, "IHDR", 4)) {
            if(len != 13) corrupt("bad IHDR");
            w = be32(body); h = be32(body + 4); depth = body[8]; color = body[9]; interlace = body[12];
            if(body[10] != 0 || body[11] != 0 || interlace > 1) corrupt("unsupported PNG method");
            header = true;
        } else if(!std::memcmp(type, "PLTE", 4)) {
            palette.assign(body, body + len);
        } else if(!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if(!std::memcmp(type, "IEND", 4)) {
            ended = true;
        } else if(!(type[0] & 32)) {
            corrupt("unknown critical PNG chunk");
        }
        pos += 12 + size_t(len);
    }
    if(!header || w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24)) corrupt("bad PNG header");
    int channels;
    switch(color) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: corrupt("bad PNG colour type");
    }
    bool depth_ok = depth == 8 || (depth == 16 && color != 3) ||
                    ((depth == 1 || depth == 2 || depth == 4) && (color == 0 || color == 3));
    if(!depth_ok) corrupt("bad PNG bit depth");
    // stb_image's size limit for PNG: at most 2^30 bytes of samples per
    // channel count, in its integer form (it rejects larger files as too large)
    if((1u << 30) / w / uint32_t(channels) < h) corrupt("PNG image too large");
    if(color == 3 && (palette.empty() || palette.size() % 3 || palette.size() > 768)) corrupt("bad PNG palette");
    if(idat.size() < 2) corrupt("missing PNG image data");
    if((idat[0] & 15) != 8 || ((idat[0] << 8) | idat[1]) % 31 || (idat[1] & 32)) corrupt("bad zlib header");
    const size_t bits_per_pixel = size_t(channels) * size_t(depth);
    const size_t bpp = std::max<size_t>(1, bits_per_pixel / 8);
    std::vector<uint8_t> raw = inflate(idat.data() + 2, idat.size() - 2, size_t(h) * (size_t(w) * bits_per_pixel / 8 + 1));

    width = int(w); height = int(h);
    std::vector<uint8_t> rgb(size_t(w) * size_t(h) * 3);
    const int scale = color == 0 ? (depth == 1 ? 0xff : depth == 2 ? 0x55 : depth == 4 ? 0x11 : 1) : 1;
    auto sample = [&](const uint8_t *row, size_t x, int c) -> int {
        // Sample c of pixel x of an unfiltered row, reduced to 8 bits.
        if(depth == 8) return row[x * size_t(channels) + size_t(c)];
        if(depth == 16) return row[(x * size_t(channels) + size_t(c)) * 2];
        size_t bit = x * size_t(depth);
        int v = (row[bit / 8] >> (8 - depth - int(bit % 8))) & ((1 << depth) - 1);
        return v;
    };
    auto store = [&](const uint8_t *row, size_t x, size_t out_x, size_t out_y) {
        uint8_t *dst = &rgb[(out_y * size_t(w) + out_x) * 3];
        if(color == 3) {
            size_t index = size_t(sample(row, x, 0));
            if(3 * index + 2 >= palette.size()) corrupt("PNG palette index out of range");
            dst[0] = palette[3 * index]; dst[1] = palette[3 * index + 1]; dst[2] = palette[3 * index + 2];
        } else if(channels <= 2) {
            uint8_t g = uint8_t(sample(row, x, 0) * scale);
            dst[0] = dst[1] = dst[2] = g;
        } else {
            for(int c = 0; c < 3; c++) dst[c] = uint8_t(sample(row, x, c));
        }
    };
    std::vector<uint8_t> rows;
    size_t offset = 0;
    if(!interlace) {
        size_t stride = (size_t(w) * bits_per_pixel + 7) / 8;
        if(raw.size() < size_t(h) * (stride + 1)) corrupt("truncated PNG image data");
        unfilter(raw.data(), h, stride, bpp, rows);
        for(size_t y = 0; y < h; y++)
            for(size_t x = 0; x < w; x++) store(&rows[y * stride], x, x, y);
    } else {
        static const int x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
        static const int dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
        for(int p = 0; p < 7; p++) {
            size_t pw = (size_t(w) - size_t(x0[p]) + size_t(dx[p]) - 1) / size_t(dx[p]);
            size_t ph = (size_t(h) - size_t(y0[p]) + size_t(dy[p]) - 1) / size_t(dy[p]);
            if(size_t(w) <= size_t(x0[p])) pw = 0;
            if(size_t(h) <= size_t(y0[p])) ph = 0;
            if(!pw || !ph) continue;
            size_t stride = (pw * bits_per_pixel + 7) / 8;
            if(raw.size() < offset + ph * (stride + 1)) corrupt("truncated PNG image data");
            unfilter(raw.data() + offset, ph, stride, bpp, rows);
            offset += ph * (stride + 1);
            for(size_t y = 0; y < ph; y++)
                for(size_t x = 0; x < pw; x++)
                    store(&rows[y * stride], x, size_t(x0[p]) + x * size_t(dx[p]), size_t(y0[p]) + y * size_t(dy[p]));
        }
    }
    return rgb;
}

// ---------------------------------------------------------------- PNM
std::vector<uint8_t> decode_pnm(const std::vector<uint8_t>& file, int& width, int& height) {
    if(file.size() < 2 || file[0] != 'P' || (file[1] != '5' && file[1] != '6')) corrupt("not a binary PNM file");
    const int channels = file[1] == '6' ? 3 : 1;
    size_t pos = 2;
    auto number = [&]() -> long {
        for(;;) {
            while(pos < file.size() && std::isspace(file[pos])) pos++;
            if(pos < file.size() && file[pos] == '#') { while(pos < file.size() && file[pos] != '\n' && file[pos] != '\r') pos++; continue; }
            break;
        }
        if(pos >= file.size() || !std::isdigit(file[pos])) corrupt("bad PNM header");
        long v = 0;
        while(pos < file.size() && std::isdigit(file[pos])) {
            v = v * 10 + (file[pos++] - '0');
            if(v > (1L << 24)) corrupt("bad PNM header");
        }
        return v;
    };
    long w = number(), h = number(), maxval = number();
    if(w <= 0 || h <= 0 || maxval <= 0 || maxval > 255) corrupt("unsupported PNM header");
    if(pos >= file.size() || !std::isspace(file[pos])) corrupt("bad PNM header");
    pos++;
    size_t n = size_t(w) * size_t(h) * size_t(channels);
    if(file.size() - pos < n) corrupt("truncated PNM data");
    width = int(w); height = int(h);
    std::vector<uint8_t> rgb(size_t(w) * size_t(h) * 3);
    for(size_t i = 0; i < size_t(w) * size_t(h); i++)
        for(int c = 0; c < 3; c++) rgb[3 * i + size_t(c)] = file[pos + i * size_t(channels) + size_t(channels == 3 ? c : 0)];
    return rgb;
}

} // namespace

// ---------------------------------------------------------------- JPEG
// Baseline and progressive Huffman JPEG (ITU-T T.81), 8-bit, 1/3/4
// components, decoded to the values stb_image (v2.30, the reference's
// vendored loader) returns: coefficients dequantized into 16-bit integers,
// the integer "islow" IDCT of IJG's jidctint.c with stb's rounding (12-bit
// constants, 2 guard bits after the column pass, +128 and a clamp at the
// end), stb's centred upsampling filters for 2x horizontal, vertical or
// both (nearest for other ratios) and its fixed-point YCbCr conversion.
namespace {

const uint8_t zigzag_to_natural[64 + 15] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
    21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60,
    61, 54, 47, 55, 62, 63,
    // runs past the end of a corrupt block land on the last coefficient
    63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct jpeg_huffman {
    uint8_t values[256] = {};
    int32_t maxcode[18] = {}, valptr[17] = {}, mincode[17] = {};
    bool defined = false;
    void build(const int counts[16]) {
        int code = 0, k = 0;
        for(int len = 1; len <= 16; len++) {
            valptr[len] = k;
            mincode[len] = code;
            code += counts[len - 1];
            k += counts[len - 1];
            if(counts[len - 1] && code - 1 >= (1 << len)) corrupt("bad Huffman code lengths");
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        defined = true;
    }
};

struct jpeg_component {
    int id = 0, h = 1, v = 1, tq = 0, hd = 0, ha = 0;
    int x = 0, y = 0, w2 = 0, h2 = 0, dc_pred = 0;
    std::vector<uint8_t> pixels;   // w2 x h2
    std::vector<int16_t> coeff;    // progressive: 64 per block, (w2/8) x (h2/8) blocks
};

class jpeg_decoder {
public:
    explicit jpeg_decoder(const std::vector<uint8_t>& f) : file(f) {}

    std::vector<uint8_t> decode(int& width, int& height) {
        if(file.size() < 2 || file[0] != 0xFF || file[1] != 0xD8) corrupt("not a JPEG file");
        pos = 2;
        // Header: table and application segments up to the frame; stray
        // bytes between segments are skipped.
        int m = next_marker();
        while(m != 0xC0 && m != 0xC1 && m != 0xC2) {
            process_marker(m);
            m = next_marker();
            while(m < 0) {
                if(pos >= file.size()) corrupt("no JPEG frame");
                m = next_marker();
            }
        }
        progressive = m == 0xC2;
        frame_header();
        // Body: scans and tables up to EOI.  As in stb_image, a marker the
        // decoder does not know (or the end of the file) ends decoding with
        // what has been decoded so far.
        m = next_marker();
        while(m != 0xD9) {
            if(m == 0xDA) {
                scan_header();
                entropy_coded_data();
                if(marker < 0) marker = marker_after_junk();
                m = marker >= 0 ? marker : next_marker();
                marker = -1;
                if(m >= 0xD0 && m <= 0xD7) m = next_marker();
            } else if(m == 0xDC) {  // DNL
                if(be16() != 4 || be16() != height_px) corrupt("bad DNL");
                m = next_marker();
            } else {
                try {
                    process_marker(m);
                } catch(const std::runtime_error&) {
                    break;
                }
                m = next_marker();
            }
        }
        if(progressive) finish_progressive();
        width = width_px; height = height_px;
        return assemble();
    }

private:
    const std::vector<uint8_t>& file;
    size_t pos = 0;
    int width_px = 0, height_px = 0, ncomp = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, jfif = false;
    int adobe_transform = -1, rgb_ids = 0;
    int restart_interval = 0;
    uint16_t dequant[4][64] = {};
    jpeg_huffman huff_dc[4], huff_ac[4];
    jpeg_component comp[4];
    int scan_n = 0, order[4] = {}, spec_start = 0, spec_end = 63, succ_high = 0, succ_low = 0;
    // entropy decoder state
    uint32_t bits = 0;
    int nbits = 0, marker = -1, eob_run = 0, todo = 0;
    bool nomore = false;

    uint8_t byte() { if(pos >= file.size()) corrupt("truncated JPEG"); return file[pos++]; }
    int be16() { int a = byte(); return (a << 8) | byte(); }
    // Skips entropy-coded junk after a scan: returns the first 0xFF-prefixed
    // byte that is neither a stuffed zero nor fill, or -1 at the end.
    int marker_after_junk() {
        while(pos < file.size()) {
            int x = file[pos++];
            while(x == 0xFF) {
                if(pos >= file.size()) return -1;
                x = file[pos++];
                if(x != 0x00 && x != 0xFF) return x;
            }
        }
        return -1;
    }
    static bool fits_int(long long v) { return v >= INT32_MIN && v <= INT32_MAX; }
    static bool fits_short(long long v) { return v >= -32768 && v <= 32767; }
    int next_marker() {
        if(pos >= file.size()) return -1;
        int x = byte();
        if(x != 0xFF) return -1;
        while(x == 0xFF) { if(pos >= file.size()) return -1; x = byte(); }
        return x;
    }

    void process_marker(int m) {
        if(m == 0xDD) {
            if(be16() != 4) corrupt("bad DRI");
            restart_interval = be16();
        } else if(m == 0xDB) {
            int L = be16() - 2;
            while(L > 0) {
                int q = byte(), p = q >> 4, t = q & 15;
                if(p > 1 || t > 3) corrupt("bad DQT");
                for(int i = 0; i < 64; i++) dequant[t][zigzag_to_natural[i]] = uint16_t(p ? be16() : byte());
                L -= p ? 129 : 65;
            }
            if(L != 0) corrupt("bad DQT length");
        } else if(m == 0xC4) {
            int L = be16() - 2;
            while(L > 0) {
                int q = byte(), tc = q >> 4, th = q & 15;
                if(tc > 1 || th > 3) corrupt("bad DHT");
                int counts[16], n = 0;
                for(int i = 0; i < 16; i++) { counts[i] = byte(); n += counts[i]; }
                if(n > 256) corrupt("bad DHT");
                jpeg_huffman& h = tc ? huff_ac[th] : huff_dc[th];
                h.build(counts);
                for(int i = 0; i < n; i++) h.values[i] = byte();
                L -= 17 + n;
            }
            if(L != 0) corrupt("bad DHT length");
        } else if((m >= 0xE0 && m <= 0xEF) || m == 0xFE) {
            int L = be16();
            if(L < 2) corrupt("bad APP length");
            L -= 2;
            const size_t start = pos;
            if(m == 0xE0 && L >= 5 && file.size() - pos >= 5 && std::memcmp(&file[pos], "JFIF\0", 5) == 0) jfif = true;
            if(m == 0xEE && L >= 12 && file.size() - pos >= 12 && std::memcmp(&file[pos], "Adobe", 6) == 0)
                adobe_transform = file[pos + 11];
            pos = start + size_t(L);
        } else {
            corrupt("unsupported JPEG marker");
        }
    }

    void frame_header() {
        const int L = be16();
        if(byte() != 8) corrupt("only 8-bit JPEG");
        height_px = be16();
        width_px = be16();
        if(height_px == 0 || width_px == 0) corrupt("bad JPEG size");
        ncomp = byte();
        if(3LL * width_px * height_px + 1 > INT32_MAX || 1LL * width_px * height_px * ncomp > INT32_MAX)
            corrupt("JPEG too large");
        if(ncomp != 1 && ncomp != 3 && ncomp != 4) corrupt("bad JPEG component count");
        if(L != 8 + 3 * ncomp) corrupt("bad SOF length");
        static const uint8_t rgb[3] = {'R', 'G', 'B'};
        for(int i = 0; i < ncomp; i++) {
            comp[i].id = byte();
            if(ncomp == 3 && comp[i].id == rgb[i]) rgb_ids++;
            int q = byte();
            comp[i].h = q >> 4; comp[i].v = q & 15;
            if(comp[i].h < 1 || comp[i].h > 4 || comp[i].v < 1 || comp[i].v > 4) corrupt("bad sampling factors");
            comp[i].tq = byte();
            if(comp[i].tq > 3) corrupt("bad quantization table");
        }
        for(int i = 0; i < ncomp; i++) { hmax = std::max(hmax, comp[i].h); vmax = std::max(vmax, comp[i].v); }
        for(int i = 0; i < ncomp; i++)
            if(hmax % comp[i].h || vmax % comp[i].v) corrupt("bad sampling factors");
        mcux = (width_px + hmax * 8 - 1) / (hmax * 8);
        mcuy = (height_px + vmax * 8 - 1) / (vmax * 8);
        for(int i = 0; i < ncomp; i++) {
            jpeg_component& c = comp[i];
            c.x = (width_px * c.h + hmax - 1) / hmax;
            c.y = (height_px * c.v + vmax - 1) / vmax;
            c.w2 = mcux * c.h * 8;
            c.h2 = mcuy * c.v * 8;
            c.pixels.assign(size_t(c.w2) * size_t(c.h2), 0);
            if(progressive) c.coeff.assign(size_t(c.w2 / 8) * size_t(c.h2 / 8) * 64, 0);
        }
    }

    void scan_header() {
        const int L = be16();
        scan_n = byte();
        if(scan_n < 1 || scan_n > 4 || scan_n > ncomp) corrupt("bad SOS component count");
        if(L != 6 + 2 * scan_n) corrupt("bad SOS length");
        for(int i = 0; i < scan_n; i++) {
            const int id = byte(), q = byte();
            int which = -1;
            for(int k = 0; k < ncomp; k++) if(comp[k].id == id) which = k;
            if(which < 0) corrupt("bad SOS component");
            comp[which].hd = q >> 4; comp[which].ha = q & 15;
            if(comp[which].hd > 3 || comp[which].ha > 3) corrupt("bad SOS table");
            order[i] = which;
        }
        spec_start = byte(); spec_end = byte();
        const int a = byte();
        succ_high = a >> 4; succ_low = a & 15;
        if(progressive) {
            if(spec_start > 63 || spec_end > 63 || spec_start > spec_end || succ_high > 13 || succ_low > 13)
                corrupt("bad progressive scan");
        } else {
            if(spec_start != 0 || succ_high != 0 || succ_low != 0) corrupt("bad baseline scan");
            spec_end = 63;
        }
    }

    // ---- entropy-coded segment: 0xFF00 is a data byte; any other marker
    // ends the segment and further bits read as zeros.
    void fill() {
        while(nbits <= 24) {
            int b = 0;
            if(!nomore) {
                b = pos < file.size() ? file[pos++] : 0;
                if(b == 0xFF) {
                    int c = pos < file.size() ? file[pos++] : 0;
                    while(c == 0xFF) c = pos < file.size() ? file[pos++] : 0;
                    if(c != 0) { marker = c; nomore = true; b = 0; }
                }
            }
            bits |= uint32_t(b) << (24 - nbits);
            nbits += 8;
        }
    }
    int get_bit() {
        if(nbits < 1) fill();
        const int b = int(bits >> 31);
        bits <<= 1; nbits--;
        return b;
    }
    int get_bits(int n) {
        if(n == 0) return 0;
        if(nbits < n) fill();
        const int v = int(bits >> (32 - n));
        bits <<= n; nbits -= n;
        return v;
    }
    int receive_extend(int n) {
        const int v = get_bits(n);
        return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v;
    }
    int decode(const jpeg_huffman& h) {
        if(!h.defined) corrupt("undefined Huffman table");
        int code = 0;
        for(int len = 1; len <= 16; len++) {
            code = (code << 1) | get_bit();
            if(h.maxcode[len] >= 0 && code <= h.maxcode[len]) {
                const int k = h.valptr[len] + code - h.mincode[len];
                if(k < 0 || k > 255) corrupt("bad Huffman code");
                return h.values[k];
            }
        }
        corrupt("bad Huffman code");
    }
    void reset() {
        bits = 0; nbits = 0; nomore = false; marker = -1; eob_run = 0;
        for(int i = 0; i < ncomp; i++) comp[i].dc_pred = 0;
        todo = restart_interval ? restart_interval : 0x7fffffff;
    }
    // After each MCU: at the end of a restart interval expect RSTn; false
    // ends the scan (anything else is left for the marker loop).
    bool restart_point() {
        if(--todo > 0) return true;
        if(nbits < 24) fill();
        if(!(marker >= 0xD0 && marker <= 0xD7)) return false;
        reset();
        return true;
    }

    void decode_block(int16_t data[64], jpeg_component& c) {
        std::memset(data, 0, 64 * sizeof(int16_t));
        const int t = decode(huff_dc[c.hd]);
        if(t > 15) corrupt("bad DC code");
        const int diff = t ? receive_extend(t) : 0;
        if(!fits_int(1LL * c.dc_pred + diff)) corrupt("bad DC delta");
        c.dc_pred += diff;
        const uint16_t *dq = dequant[c.tq];
        if(!fits_short(1LL * c.dc_pred * dq[0])) corrupt("DC coefficient overflow");
        data[0] = int16_t(c.dc_pred * int(dq[0]));
        int k = 1;
        do {
            const int rs = decode(huff_ac[c.ha]);
            const int s = rs & 15, r = rs >> 4;
            if(s == 0) {
                if(rs != 0xF0) break;
                k += 16;
            } else {
                k += r;
                const int zig = zigzag_to_natural[k++];
                data[zig] = int16_t(receive_extend(s) * int(dq[zig]));
            }
        } while(k < 64);
    }

    void decode_dc_progressive(int16_t data[64], jpeg_component& c) {
        if(spec_end != 0) corrupt("bad progressive DC scan");
        if(succ_high == 0) {
            std::memset(data, 0, 64 * sizeof(int16_t));
            const int t = decode(huff_dc[c.hd]);
            if(t > 15) corrupt("bad DC code");
            const int diff = t ? receive_extend(t) : 0;
            if(!fits_int(1LL * c.dc_pred + diff)) corrupt("bad DC delta");
            c.dc_pred += diff;
            if(!fits_short(1LL * c.dc_pred * (1 << succ_low))) corrupt("DC coefficient overflow");
            data[0] = int16_t(c.dc_pred * (1 << succ_low));
        } else if(get_bit()) {
            data[0] = int16_t(data[0] + (1 << succ_low));
        }
    }

    void decode_ac_progressive(int16_t data[64], jpeg_component& c) {
        if(spec_start == 0) corrupt("bad progressive AC scan");
        if(succ_high == 0) {
            if(eob_run) { eob_run--; return; }
            int k = spec_start;
            do {
                const int rs = decode(huff_ac[c.ha]);
                const int s = rs & 15, r = rs >> 4;
                if(s == 0) {
                    if(r < 15) {
                        eob_run = (1 << r) - 1;
                        if(r) eob_run += get_bits(r);
                        break;
                    }
                    k += 16;
                } else {
                    k += r;
                    const int zig = zigzag_to_natural[k++];
                    data[zig] = int16_t(receive_extend(s) * (1 << succ_low));
                }
            } while(k <= spec_end);
            return;
        }
        const int16_t bit = int16_t(1 << succ_low);
        auto refine = [&](int16_t *p) {
            if(get_bit() && (*p & bit) == 0) *p = int16_t(*p > 0 ? *p + bit : *p - bit);
        };
        if(eob_run) {
            eob_run--;
            for(int k = spec_start; k <= spec_end; k++) {
                int16_t *p = &data[zigzag_to_natural[k]];
                if(*p != 0) refine(p);
            }
            return;
        }
        int k = spec_start;
        do {
            const int rs = decode(huff_ac[c.ha]);
            int s = rs & 15, r = rs >> 4;
            if(s == 0) {
                if(r < 15) {
                    eob_run = (1 << r) - 1;
                    if(r) eob_run += get_bits(r);
                    r = 64;  // refine the rest of the band, then end the block
                }
            } else {
                if(s != 1) corrupt("bad refinement code");
                s = get_bit() ? bit : -bit;
            }
            while(k <= spec_end) {
                int16_t *p = &data[zigzag_to_natural[k++]];
                if(*p != 0) refine(p);
                else {
                    if(r == 0) { *p = int16_t(s); break; }
                    --r;
                }
            }
        } while(k <= spec_end);
    }

    void entropy_coded_data() {
        reset();
        int16_t block[64];
        if(scan_n == 1) {
            jpeg_component& c = comp[order[0]];
            const int w = (c.x + 7) >> 3, h = (c.y + 7) >> 3;
            for(int j = 0; j < h; j++)
                for(int i = 0; i < w; i++) {
                    if(!progressive) {
                        decode_block(block, c);
                        idct(block, &c.pixels[size_t(c.w2) * size_t(j * 8) + size_t(i * 8)], c.w2);
                    } else {
                        int16_t *data = &c.coeff[64 * (size_t(i) + size_t(j) * size_t(c.w2 / 8))];
                        if(spec_start == 0) decode_dc_progressive(data, c);
                        else decode_ac_progressive(data, c);
                    }
                    if(!restart_point()) return;
                }
            return;
        }
        if(progressive && spec_start != 0) corrupt("interleaved progressive AC scan");
        for(int j = 0; j < mcuy; j++)
            for(int i = 0; i < mcux; i++) {
                for(int k = 0; k < scan_n; k++) {
                    jpeg_component& c = comp[order[k]];
                    for(int y = 0; y < c.v; y++)
                        for(int x = 0; x < c.h; x++) {
                            const int bx = i * c.h + x, by = j * c.v + y;
                            if(!progressive) {
                                decode_block(block, c);
                                idct(block, &c.pixels[size_t(c.w2) * size_t(by * 8) + size_t(bx * 8)], c.w2);
                            } else {
                                decode_dc_progressive(&c.coeff[64 * (size_t(bx) + size_t(by) * size_t(c.w2 / 8))], c);
                            }
                        }
                }
                if(!restart_point()) return;
            }
    }

    void finish_progressive() {
        for(int n = 0; n < ncomp; n++) {
            jpeg_component& c = comp[n];
            const int w = (c.x + 7) >> 3, h = (c.y + 7) >> 3;
            for(int j = 0; j < h; j++)
                for(int i = 0; i < w; i++) {
                    int16_t *data = &c.coeff[64 * (size_t(i) + size_t(j) * size_t(c.w2 / 8))];
                    for(int k = 0; k < 64; k++) data[k] = int16_t(data[k] * int(dequant[c.tq][k]));
                    idct(data, &c.pixels[size_t(c.w2) * size_t(j * 8) + size_t(i * 8)], c.w2);
                }
        }
    }

    // Integer IDCT (jidctint "islow" with 12-bit constants; stb's rounding).
    // The sums and products wrap modulo 2^32, as stb's int arithmetic does on
    // the targets it is built for, but without undefined behaviour: extreme
    // coefficients of a malformed file can overflow 32 bits.
    static int fix(float x) { return int(double(x * 4096) + 0.5); }
    static uint8_t clamp255(int x) { return uint8_t(x < 0 ? 0 : x > 255 ? 255 : x); }
    using w32 = uint32_t;
    static int shifted(w32 x, int bits) { return int32_t(x) >> bits; }
    struct idct_terms { w32 t0, t1, t2, t3, x0, x1, x2, x3; };
    static idct_terms idct_1d(int s0, int s1, int s2, int s3, int s4, int s5, int s6, int s7) {
        static const w32 c0541 = w32(fix(0.5411961f)), cm1847 = w32(fix(-1.847759065f)), c0765 = w32(fix(0.765366865f));
        static const w32 c1175 = w32(fix(1.175875602f)), c0298 = w32(fix(0.298631336f)), c2053 = w32(fix(2.053119869f));
        static const w32 c3072 = w32(fix(3.072711026f)), c1501 = w32(fix(1.501321110f)), cm0899 = w32(fix(-0.899976223f));
        static const w32 cm2562 = w32(fix(-2.562915447f)), cm1961 = w32(fix(-1.961570560f)), cm0390 = w32(fix(-0.390180644f));
        idct_terms o;
        w32 p2 = w32(s2), p3 = w32(s6);
        w32 p1 = (p2 + p3) * c0541;
        w32 t2 = p1 + p3 * cm1847, t3 = p1 + p2 * c0765;
        p2 = w32(s0); p3 = w32(s4);
        w32 t0 = (p2 + p3) * 4096u, t1 = (p2 - p3) * 4096u;
        o.x0 = t0 + t3; o.x3 = t0 - t3; o.x1 = t1 + t2; o.x2 = t1 - t2;
        t0 = w32(s7); t1 = w32(s5); t2 = w32(s3); t3 = w32(s1);
        p3 = t0 + t2;
        w32 p4 = t1 + t3;
        p1 = t0 + t3;
        p2 = t1 + t2;
        const w32 p5 = (p3 + p4) * c1175;
        t0 = t0 * c0298; t1 = t1 * c2053; t2 = t2 * c3072; t3 = t3 * c1501;
        p1 = p5 + p1 * cm0899; p2 = p5 + p2 * cm2562;
        p3 = p3 * cm1961; p4 = p4 * cm0390;
        t3 += p1 + p4; t2 += p2 + p3; t1 += p2 + p4; t0 += p1 + p3;
        o.t0 = t0; o.t1 = t1; o.t2 = t2; o.t3 = t3;
        return o;
    }
    static void idct(const int16_t *d, uint8_t *out, int stride) {
        int v[64];
        for(int i = 0; i < 8; i++) {
            const int16_t *c = d + i;
            if(!c[8] && !c[16] && !c[24] && !c[32] && !c[40] && !c[48] && !c[56]) {
                const int dc = c[0] * 4;
                for(int r = 0; r < 8; r++) v[i + 8 * r] = dc;
                continue;
            }
            idct_terms o = idct_1d(c[0], c[8], c[16], c[24], c[32], c[40], c[48], c[56]);
            o.x0 += 512; o.x1 += 512; o.x2 += 512; o.x3 += 512;
            v[i] = shifted(o.x0 + o.t3, 10); v[i + 56] = shifted(o.x0 - o.t3, 10);
            v[i + 8] = shifted(o.x1 + o.t2, 10); v[i + 48] = shifted(o.x1 - o.t2, 10);
            v[i + 16] = shifted(o.x2 + o.t1, 10); v[i + 40] = shifted(o.x2 - o.t1, 10);
            v[i + 24] = shifted(o.x3 + o.t0, 10); v[i + 32] = shifted(o.x3 - o.t0, 10);
        }
        for(int r = 0; r < 8; r++) {
            const int *w = v + 8 * r;
            idct_terms o = idct_1d(w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7]);
            const w32 bias = 65536 + (128 << 17);
            o.x0 += bias; o.x1 += bias; o.x2 += bias; o.x3 += bias;
            uint8_t *p = out + size_t(r) * size_t(stride);
            p[0] = clamp255(shifted(o.x0 + o.t3, 17)); p[7] = clamp255(shifted(o.x0 - o.t3, 17));
            p[1] = clamp255(shifted(o.x1 + o.t2, 17)); p[6] = clamp255(shifted(o.x1 - o.t2, 17));
            p[2] = clamp255(shifted(o.x2 + o.t1, 17)); p[5] = clamp255(shifted(o.x2 - o.t1, 17));
            p[3] = clamp255(shifted(o.x3 + o.t0, 17)); p[4] = clamp255(shifted(o.x3 - o.t0, 17));
        }
    }

    // ---- upsampling (centred triangle filters for factors of 2) and colour
    static void upsample_row(const uint8_t *near_row, const uint8_t *far_row, int w, int hs, int vs, uint8_t *out) {
        if(hs == 1 && vs == 1) { std::memcpy(out, near_row, size_t(w)); return; }
        if(hs == 1 && vs == 2) {
            for(int i = 0; i < w; i++) out[i] = uint8_t((3 * near_row[i] + far_row[i] + 2) >> 2);
            return;
        }
        if(hs == 2 && vs == 1) {
            if(w == 1) { out[0] = out[1] = near_row[0]; return; }
            out[0] = near_row[0];
            out[1] = uint8_t((near_row[0] * 3 + near_row[1] + 2) >> 2);
            int i = 1;
            for(; i < w - 1; i++) {
                const int n = 3 * near_row[i] + 2;
                out[2 * i] = uint8_t((n + near_row[i - 1]) >> 2);
                out[2 * i + 1] = uint8_t((n + near_row[i + 1]) >> 2);
            }
            out[2 * i] = uint8_t((near_row[w - 2] * 3 + near_row[w - 1] + 2) >> 2);
            out[2 * i + 1] = near_row[w - 1];
            return;
        }
        if(hs == 2 && vs == 2) {
            if(w == 1) { out[0] = out[1] = uint8_t((3 * near_row[0] + far_row[0] + 2) >> 2); return; }
            int t1 = 3 * near_row[0] + far_row[0];
            out[0] = uint8_t((t1 + 2) >> 2);
            for(int i = 1; i < w; i++) {
                const int t0 = t1;
                t1 = 3 * near_row[i] + far_row[i];
                out[2 * i - 1] = uint8_t((3 * t0 + t1 + 8) >> 4);
                out[2 * i] = uint8_t((3 * t1 + t0 + 8) >> 4);
            }
            out[2 * w - 1] = uint8_t((t1 + 2) >> 2);
            return;
        }
        for(int i = 0; i < w; i++)
            for(int j = 0; j < hs; j++) out[i * hs + j] = near_row[i];
    }

    static uint8_t blinn(int x, int y) { const unsigned t = unsigned(x * y + 128); return uint8_t((t + (t >> 8)) >> 8); }

    std::vector<uint8_t> assemble() {
        const bool is_rgb = ncomp == 3 && (rgb_ids == 3 || (adobe_transform == 0 && !jfif));
        const int decode_n = ncomp;
        struct row_state { int hs, vs, ystep, w_lores, ypos; size_t line0, line1; std::vector<uint8_t> buf; };
        std::vector<row_state> rs(static_cast<size_t>(decode_n));
        for(int k = 0; k < decode_n; k++) {
            row_state& r = rs[size_t(k)];
            r.hs = hmax / comp[k].h; r.vs = vmax / comp[k].v;
            r.ystep = r.vs >> 1;
            r.w_lores = (width_px + r.hs - 1) / r.hs;
            r.ypos = 0; r.line0 = r.line1 = 0;
            r.buf.assign(size_t(width_px) + 3, 0);
        }
        auto fixed = [](float x) { return int(x * 4096.0f + 0.5f) << 8; };
        const int cr_r = fixed(1.40200f), cr_g = -fixed(0.71414f), cb_g = -fixed(0.34414f), cb_b = fixed(1.77200f);
        std::vector<uint8_t> rgb(size_t(width_px) * size_t(height_px) * 3);
        std::vector<const uint8_t *> row(static_cast<size_t>(decode_n));
        for(int j = 0; j < height_px; j++) {
            for(int k = 0; k < decode_n; k++) {
                row_state& r = rs[size_t(k)];
                const bool bottom = r.ystep >= (r.vs >> 1);
                const uint8_t *a = &comp[k].pixels[bottom ? r.line1 : r.line0];
                const uint8_t *b = &comp[k].pixels[bottom ? r.line0 : r.line1];
                if(r.hs == 1 && r.vs == 1) row[size_t(k)] = a;
                else { upsample_row(a, b, r.w_lores, r.hs, r.vs, r.buf.data()); row[size_t(k)] = r.buf.data(); }
                if(++r.ystep >= r.vs) {
                    r.ystep = 0;
                    r.line0 = r.line1;
                    if(++r.ypos < comp[k].y) r.line1 += size_t(comp[k].w2);
                }
            }
            uint8_t *out = &rgb[size_t(j) * size_t(width_px) * 3];
            const uint8_t *y = row[0];
            for(int i = 0; i < width_px; i++, out += 3) {
                if(ncomp == 1) { out[0] = out[1] = out[2] = y[i]; continue; }
                if(ncomp == 3 && is_rgb) { out[0] = y[i]; out[1] = row[1][i]; out[2] = row[2][i]; continue; }
                if(ncomp == 4 && adobe_transform == 0) {  // CMYK
                    const uint8_t m = row[3][i];
                    out[0] = blinn(y[i], m); out[1] = blinn(row[1][i], m); out[2] = blinn(row[2][i], m);
                    continue;
                }
                const int y_fixed = (y[i] << 20) + (1 << 19);
                const int cr = row[2][i] - 128, cb = row[1][i] - 128;
                int r = y_fixed + cr * cr_r;
                int g = y_fixed + (cr * cr_g) + int(unsigned(cb * cb_g) & 0xffff0000u);
                int b = y_fixed + cb * cb_b;
                r >>= 20; g >>= 20; b >>= 20;
                out[0] = clamp255(r); out[1] = clamp255(g); out[2] = clamp255(b);
                if(ncomp == 4 && adobe_transform == 2) {  // YCCK
                    const uint8_t m = row[3][i];
                    out[0] = blinn(255 - out[0], m); out[1] = blinn(255 - out[1], m); out[2] = blinn(255 - out[2], m);
                }
            }
        }
        return rgb;
    }
};

} // namespace

namespace pts {

// Decode an image file to 8-bit RGB; throws "could not load image: <path>".
std::vector<uint8_t> load_image(const std::string& path, int& width, int& height) {
    std::ifstream in(path, std::ios::binary);
    std::vector<uint8_t> file;
    if(in) file.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    try {
        if(!in.good() && !in.eof()) throw std::runtime_error("unreadable");
        if(file.size() >= 8 && file[0] == 137 && file[1] == 'P' && file[2] == 'N' && file[3] == 'G')
            return decode_png(file, width, height);
        if(file.size() >= 2 && file[0] == 'P' && (file[1] == '5' || file[1] == '6'))
            return decode_pnm(file, width, height);
        if(file.size() >= 2 && file[0] == 0xFF && file[1] == 0xD8)
            return jpeg_decoder(file).decode(width, height);
        throw std::runtime_error("unsupported image format");
    } catch(const std::runtime_error&) {
        throw std::runtime_error("could not load image: " + path);
    }
}

// The reference's color_channel (src/display/color.h:6-11).
int color_channel(float linear) {
    if(!(linear > 0)) return 0;
    if(linear >= 1) return 255;
    float gamma_corrected = std::sqrt(linear);
    return int(255.99 * gamma_corrected);
}

} // namespace pts

extern "C" {

int32_t pts_load_image(const char *path, int32_t *width, int32_t *height, uint8_t **rgb) {
    try {
        if(!path || !width || !height || !rgb) throw std::invalid_argument("null argument");
        int w = 0, h = 0;
        std::vector<uint8_t> pixels = pts::load_image(path, w, h);
        auto *out = static_cast<uint8_t *>(std::malloc(pixels.size()));
        if(!out) throw std::runtime_error("out of memory");
        std::memcpy(out, pixels.data(), pixels.size());
        *width = w; *height = h; *rgb = out;
        return 0;
    } catch(const std::exception& e) { pts::set_error(e.what()); return -1; }
}

void pts_free_image(uint8_t *rgb) { std::free(rgb); }

int32_t pts_write_ppm(const char *path, const float *rgb, int32_t width, int32_t height) {
    // src/app/main.cpp:40-46: "P3", rows from the top of the image (the
    // accumulation buffer's row 0 is the bottom), one pixel per line.
    try {
        if(!path || !rgb || width <= 0 || height <= 0) throw std::invalid_argument("invalid PPM arguments");
        std::string text = "P3\n" + std::to_string(width) + " " + std::to_string(height) + "\n255\n";
        text.reserve(text.size() + size_t(width) * size_t(height) * 12);
        char line[64];
        for(int32_t y = height - 1; y >= 0; y--) {
            for(int32_t x = 0; x < width; x++) {
                const float *p = rgb + 3 * (size_t(y) * size_t(width) + size_t(x));
                int n = std::snprintf(line, sizeof line, "%d %d %d\n", pts::color_channel(p[0]),
                                      pts::color_channel(p[1]), pts::color_channel(p[2]));
                text.append(line, size_t(n));
            }
        }
        const bool to_stdout = std::strcmp(path, "-") == 0;
        FILE *f = to_stdout ? stdout : std::fopen(path, "wb");
        if(!f) throw std::runtime_error(std::string("could not open ") + path);
        size_t written = std::fwrite(text.data(), 1, text.size(), f);
        int flushed = to_stdout ? std::fflush(f) : std::fclose(f);
        if(written != text.size() || flushed != 0) throw std::runtime_error("could not write image output");
        return 0;
    } catch(const std::exception& e) { pts::set_error(e.what()); return -1; }
}

}
