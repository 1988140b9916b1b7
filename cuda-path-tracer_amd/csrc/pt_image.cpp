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
//   binary PNM:     P5 (grey, replicated) and P6, maxval <= 255.
// JPEG is not decoded here; such a texture fails the load with the
// reference's own "could not load image" error.
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
    out.reserve(expected);
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
