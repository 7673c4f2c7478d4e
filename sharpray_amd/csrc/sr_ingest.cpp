// Host-side ingest of the reference's asset formats, layered on the public C ABI:
//  - Wavefront OBJ + MTL (ObjLoader, obj.cs:19-512): lexer, faces fanned from their first vertex,
//    default UVs, V flip, negative indices, smooth normals for files without vn, one object per
//    o / g / usemtl block; MTL Kd / Ke / Pr / Pm / Ns / Ni / d / Tr and map_* textures (TGA).
//  - Radiance RGBE .hdr (ImageHdr.LoadHdr, image.cs:307-397) with RLE scanlines, decoded like
//    Texture.FromHdr (texture.cs:154-163) and Color.FromPixel(PixelHdr) (math.cs:196-202).
//  - TGA (Image.LoadTga, image.cs:217-266), decoded like Texture.FromSrgb / FromLinear
//    (texture.cs:138-152, math.cs:182-194).
// Errors carry the reference's exception texts. Inputs the reference would loop on forever (an RLE
// code of 0) or index out of range with are errors here.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "../../include/sharpray.h"
#include "sr_host.h"

namespace {

using sr::sr_set_error;

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    uint8_t buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

// File.Exists: true for regular files only.
bool file_exists(const std::string& path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// Path.GetDirectoryName / Path.Combine / Path.GetExtension for '/' paths.
std::string dir_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k == 0 ? 1 : k);
}
std::string combine(const std::string& dir, const std::string& name) {
    if (!name.empty() && name[0] == '/') return name;
    if (dir.empty()) return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}
std::string ext_of(const std::string& p) {
    const size_t k = p.find_last_of('/');
    const std::string f = k == std::string::npos ? p : p.substr(k + 1);
    const size_t d = f.find_last_of('.');
    return d == std::string::npos ? std::string() : f.substr(d);
}

struct IngestError {
    std::string msg;
};
[[noreturn]] void raise(const std::string& m) { throw IngestError{m}; }

// ---- ObjLexer (obj.cs:19-125) -------------------------------------------------------------------
enum class Tok { Word, Slash, Newline, End };

class Lexer {
public:
    Lexer(const uint8_t* p, size_t n) : p_(p), e_(p + n) {}

    Tok peek() {
        if (!hasPeek_) {
            peekKind_ = read(peekWord_, 512);
            hasPeek_ = true;
        }
        return peekKind_;
    }
    // Next into a caller buffer of `cap` characters (a longer word overflows it in the reference).
    Tok next(std::string& w, size_t cap) {
        if (hasPeek_) {
            hasPeek_ = false;
            if (peekWord_.size() > cap) raise("OBJ: word too long");
            w = peekWord_;
            return peekKind_;
        }
        return read(w, cap);
    }
    Tok next() {
        if (hasPeek_) {
            hasPeek_ = false;
            return peekKind_;
        }
        std::string scratch;
        return read(scratch, 64);
    }
    void skip_line() {
        Tok k;
        do {
            k = next();
        } while (k != Tok::Newline && k != Tok::End);
    }

private:
    static bool word_end(uint8_t c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '/'; }
    Tok read(std::string& w, size_t cap) {
        w.clear();
        while (true) {
            if (p_ == e_) return Tok::End;
            const uint8_t ch = *p_++;
            switch (ch) {
                case '\n': return Tok::Newline;
                case '/': return Tok::Slash;
                case ' ':
                case '\t':
                case '\r': continue;
                case '#':
                    while (true) {
                        if (p_ == e_) return Tok::End;
                        if (*p_++ == '\n') return Tok::Newline;
                    }
                default:
                    w.push_back((char)ch);
                    while (p_ != e_ && !word_end(*p_)) w.push_back((char)*p_++);
                    if (w.size() > cap) raise("OBJ: word too long");
                    return Tok::Word;
            }
        }
    }
    const uint8_t* p_;
    const uint8_t* e_;
    bool hasPeek_ = false;
    Tok peekKind_ = Tok::End;
    std::string peekWord_;
};

const char* tok_name(Tok k) {
    switch (k) {
        case Tok::Word: return "Word";
        case Tok::Slash: return "Slash";
        case Tok::Newline: return "Newline";
        default: return "End";
    }
}

// float.TryParse(NumberStyles.Float, InvariantCulture): [sign] digits [. digits] [e [sign] digits],
// or Infinity / NaN; correctly rounded (strtof), overflow to +-Infinity.
bool parse_float(const std::string& s, float& out) {
    size_t i = 0;
    const size_t n = s.size();
    if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
    std::string rest = s.substr(i);
    for (char& c : rest) c = (char)tolower((unsigned char)c);
    if (rest == "infinity" || rest == "nan") {
        const bool neg = i > 0 && s[0] == '-';
        out = rest == "nan" ? NAN : (neg ? -INFINITY : INFINITY);
        return true;
    }
    size_t digits = 0;
    while (i < n && isdigit((unsigned char)s[i])) ++i, ++digits;
    if (i < n && s[i] == '.') {
        ++i;
        while (i < n && isdigit((unsigned char)s[i])) ++i, ++digits;
    }
    if (digits == 0) return false;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
        size_t ed = 0;
        while (i < n && isdigit((unsigned char)s[i])) ++i, ++ed;
        if (ed == 0) return false;
    }
    if (i != n) return false;
    errno = 0;
    out = strtof(s.c_str(), nullptr);
    return true;
}

// int.TryParse: [sign] digits, int32 range.
bool parse_int(const std::string& s, int& out) {
    size_t i = 0;
    bool neg = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i == s.size()) return false;
    int64_t v = 0;
    for (; i < s.size(); ++i) {
        if (!isdigit((unsigned char)s[i])) return false;
        v = v * 10 + (s[i] - '0');
        if (v > 2147483648LL) return false;
    }
    if (neg) v = -v;
    if (v > 2147483647LL || v < -2147483648LL) return false;
    out = (int)v;
    return true;
}

float read_float(Lexer& lx) {
    std::string w;
    const Tok k = lx.next(w, 32);
    if (k != Tok::Word) raise(std::string("OBJ: expected number, got ") + tok_name(k));
    float v;
    if (!parse_float(w, v)) raise("OBJ: invalid number '" + w + "'");
    return v;
}

int read_int(Lexer& lx) {
    std::string w;
    const Tok k = lx.next(w, 16);
    if (k != Tok::Word) raise(std::string("OBJ: expected int, got ") + tok_name(k));
    int v;
    if (!parse_int(w, v)) raise("OBJ: invalid int '" + w + "'");
    return v;
}

int read_index(Lexer& lx, int count) {
    int idx = read_int(lx);
    idx = idx < 0 ? count + idx : idx - 1;
    if (idx < 0 || idx >= count)
        raise("OBJ: index out of bounds (" + std::to_string(idx) + ", count: " + std::to_string(count) + ")");
    return idx;
}

// ---- TGA and RGBE decoding ---------------------------------------------------------------------
struct Reader {
    const uint8_t* p;
    size_t n, at = 0;
    uint8_t byte(const char* what) {
        if (at >= n) raise(std::string(what) + ": unexpected end of file");
        return p[at++];
    }
};

// Image.LoadTga (image.cs:217-266): 24-bit TrueColor or 8-bit grayscale, uncompressed; rows are
// flipped to top-down unless the descriptor's top-left bit is set. out: RGB bytes.
void decode_tga(const uint8_t* data, size_t size, std::vector<uint8_t>& rgb, uint32_t& w, uint32_t& h) {
    Reader r{data, size};
    if (size < 18) raise("TGA: unexpected end of file");
    r.byte("TGA");
    const uint8_t colorMapType = r.byte("TGA");
    const uint8_t imageType = r.byte("TGA");
    r.at += 9;
    w = (uint32_t)data[12] | ((uint32_t)data[13] << 8);
    h = (uint32_t)data[14] | ((uint32_t)data[15] << 8);
    const uint8_t bpp = data[16], desc = data[17];
    r.at = 18;
    if (colorMapType != 0) raise("TGA: Color-mapped images not supported");
    if (imageType != 2 && imageType != 3)
        raise("TGA: Unsupported image type " + std::to_string(imageType) +
              " (only uncompressed TrueColor and Grayscale supported)");
    if (imageType == 2 && bpp != 24)
        raise("TGA: Only 24-bit TrueColor images supported, got " + std::to_string(bpp) + "-bit");
    if (imageType == 3 && bpp != 8)
        raise("TGA: Only 8-bit Grayscale images supported, got " + std::to_string(bpp) + "-bit");
    if (w == 0 || h == 0) raise("TGA: empty image");
    const size_t n = (size_t)w * h;
    rgb.resize(n * 3);
    for (size_t i = 0; i < n; ++i) {
        if (imageType == 3) {
            const uint8_t v = r.byte("TGA");
            rgb[i * 3] = rgb[i * 3 + 1] = rgb[i * 3 + 2] = v;
        } else {
            const uint8_t b = r.byte("TGA"), g = r.byte("TGA"), rr = r.byte("TGA");
            rgb[i * 3] = rr;
            rgb[i * 3 + 1] = g;
            rgb[i * 3 + 2] = b;
        }
    }
    if ((desc & 0x20) == 0)
        for (uint32_t y = 0; y < h / 2; ++y)
            for (size_t k = 0; k < (size_t)w * 3; ++k)
                std::swap(rgb[(size_t)y * w * 3 + k], rgb[(size_t)(h - 1 - y) * w * 3 + k]);
}

std::string read_line(Reader& r) {
    std::string line;
    while (true) {
        const uint8_t v = r.byte("ImageHdr");
        if (v == '\n') return line;
        if (v == '\r') continue;
        if (line.size() >= 256) raise("ImageHdr: header line longer than 256 characters");
        line.push_back((char)v);
    }
}

bool parse_uint(const std::string& s, uint32_t& out) {
    if (s.empty()) return false;
    uint64_t v = 0;
    for (char c : s) {
        if (!isdigit((unsigned char)c)) return false;
        v = v * 10 + (uint64_t)(c - '0');
        if (v > 0xFFFFFFFFull) return false;
    }
    out = (uint32_t)v;
    return true;
}

// ImageHdr.LoadHdr (image.cs:307-382): RGBE bytes per pixel, top row first.
void decode_hdr(const uint8_t* data, size_t size, std::vector<uint8_t>& rgbe, uint32_t& w, uint32_t& h) {
    Reader r{data, size};
    const std::string magic = read_line(r);
    if (magic.rfind("#?RADIANCE", 0) != 0 && magic.rfind("#?RGBE", 0) != 0)
        raise("ImageHdr: Invalid magic '" + magic + "'");
    while (!read_line(r).empty()) {
    }
    const std::string sizeLine = read_line(r);
    std::vector<std::string> parts;
    {
        size_t a = 0;
        while (true) {
            const size_t b = sizeLine.find(' ', a);
            parts.push_back(sizeLine.substr(a, b == std::string::npos ? std::string::npos : b - a));
            if (b == std::string::npos) break;
            a = b + 1;
        }
    }
    if (parts.size() != 4 || parts[0] != "-Y" || parts[2] != "+X")
        raise("ImageHdr: Unsupported size/orientation '" + sizeLine + "'");
    if (!parse_uint(parts[1], h) || !parse_uint(parts[3], w)) raise("ImageHdr: invalid size '" + sizeLine + "'");
    if ((uint64_t)w * h > (1ull << 31)) raise("ImageHdr: image too large");
    rgbe.assign((size_t)w * h * 4, 0);
    std::vector<uint8_t> line((size_t)w * 4);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t hd[4];
        for (int k = 0; k < 4; ++k) hd[k] = r.byte("ImageHdr");
        if (hd[0] == 2 && hd[1] == 2 && (hd[2] & 0x80) == 0) {
            const uint32_t lineWidth = ((uint32_t)hd[2] << 8) | hd[3];
            if (lineWidth != w)
                raise("ImageHdr: Scanline width mismatch (" + std::to_string(lineWidth) + " vs " + std::to_string(w) +
                      ")");
            for (int c = 0; c < 4; ++c) {
                for (uint32_t x = 0; x < w;) {
                    const uint8_t code = r.byte("ImageHdr");
                    if (code > 128) {
                        const uint32_t count = code - 128u;
                        const uint8_t v = r.byte("ImageHdr");
                        if (x + count > w) raise("ImageHdr: RLE run past the end of the scanline");
                        for (uint32_t i = 0; i < count; ++i) line[(size_t)(x++) * 4 + c] = v;
                    } else {
                        if (code == 0) raise("ImageHdr: RLE literal of length 0");
                        if (x + code > w) raise("ImageHdr: RLE literal past the end of the scanline");
                        for (uint32_t i = 0; i < code; ++i) line[(size_t)(x++) * 4 + c] = r.byte("ImageHdr");
                    }
                }
            }
        } else {
            if (hd[0] == 1 && hd[1] == 1 && hd[2] == 1) raise("ImageHdr: Old RLE format not supported");
            // The header is the scanline's first pixel; the reference's copy into a scanline of
            // fewer than 4 bytes throws (image.cs:371), i.e. for width 0.
            if (w == 0) raise("ImageHdr: flat scanline shorter than its first pixel (width 0)");
            memcpy(line.data(), hd, 4);
            for (size_t k = 4; k < line.size(); ++k) line[k] = r.byte("ImageHdr");
        }
        if (!line.empty()) memcpy(rgbe.data() + (size_t)y * w * 4, line.data(), line.size());
    }
}

// Texture.FromHdr (texture.cs:154-163): Color.FromPixel(PixelHdr) then ClampLuminance.
void rgbe_to_float(const std::vector<uint8_t>& rgbe, float maxLum, float* out) {
    const size_t n = rgbe.size() / 4;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* p = rgbe.data() + i * 4;
        float r = 0.0f, g = 0.0f, b = 0.0f;
        if (p[3] != 0) {
            const float scale = ldexpf(1.0f, (int)p[3] - 128) / 256.0f;
            r = (float)p[0] * scale;
            g = (float)p[1] * scale;
            b = (float)p[2] * scale;
        }
        const float lum = 0.2126f * r + 0.7152f * g + 0.0722f * b;  // Color.Luminance (math.cs:115)
        if (lum > maxLum) {
            const float s = maxLum / lum;
            r *= s;
            g *= s;
            b *= s;
        }
        out[i * 3] = r;
        out[i * 3 + 1] = g;
        out[i * 3 + 2] = b;
    }
}

// Image.Load (image.cs:204-215) of a texture file + Texture.FromSrgb / FromLinear.
int32_t load_texture(sr_ctx* ctx, const std::string& path, bool srgb) {
    const std::string ext = ext_of(path);
    if (ext != ".tga" && ext != ".ppm" && ext != ".bmp") raise("Image: Unsupported file format '" + ext + "'");
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) raise("Image: cannot read '" + path + "'");
    if (ext != ".tga") raise(std::string("Image: Loading not supported for format '") + (ext == ".ppm" ? "Ppm" : "Bmp") + "'");
    std::vector<uint8_t> rgb;
    uint32_t w, h;
    decode_tga(bytes.data(), bytes.size(), rgb, w, h);
    std::vector<float> tex(rgb.size());
    for (size_t i = 0; i < rgb.size(); ++i)
        tex[i] = srgb ? powf((float)rgb[i] / 255.0f, 2.233333333f) : (float)rgb[i] / 255.0f;
    int32_t id = -1;
    if (sr_texture_create(ctx, tex.data(), w, h, 0, 1.0f, 1.0f, &id) != SR_OK) raise(sr_last_error());
    return id;
}

// ObjLoader.ParseMtl (obj.cs:260-409).
void parse_mtl(sr_ctx* ctx, const std::string& path, std::map<std::string, sr_material>& out) {
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) raise("OBJ: cannot read '" + path + "'");
    const std::string dir = dir_of(path);
    bool have = false;
    std::string name;
    sr_material m;
    float roughness = 1.0f, ns = 0.0f;
    bool hasPr = false;
    auto reset = [&] {
        m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
        m.radiance[0] = m.radiance[1] = m.radiance[2] = 0.0f;
        roughness = 1.0f;
        m.metallic = 0.0f;
        m.transparency = 0.0f;
        m.ior = 1.5f;
        hasPr = false;
        ns = 0.0f;
        m.color_texture = m.roughness_texture = m.metallic_texture = m.normal_texture = -1;
    };
    auto build = [&] {
        if (!have) return;
        float r = hasPr ? roughness : sqrtf(2.0f / (ns + 2.0f));
        r = r < 0.0f ? 0.0f : (r > 1.0f ? 1.0f : r);  // Math.Clamp (NaN stays NaN)
        sr_material mm = m;
        mm.roughness = r;
        out[name] = mm;
    };
    reset();
    Lexer lx(bytes.data(), bytes.size());
    std::string w;
    while (true) {
        const Tok k = lx.next(w, 512);
        if (k == Tok::End) break;
        if (k != Tok::Word) continue;
        auto color3 = [&](float* c) {
            c[0] = read_float(lx);
            c[1] = read_float(lx);
            c[2] = read_float(lx);
        };
        auto map = [&](int32_t& slot, bool srgb) {
            std::string f;
            lx.next(f, 512);
            slot = load_texture(ctx, combine(dir, f), srgb);
        };
        if (w == "newmtl") {
            build();
            lx.next(name, 512);
            have = true;
            reset();
        } else if (w == "Kd") {
            color3(m.albedo);
        } else if (w == "Ke") {
            color3(m.radiance);
        } else if (w == "Pr") {
            roughness = read_float(lx);
            hasPr = true;
        } else if (w == "Pm") {
            m.metallic = read_float(lx);
        } else if (w == "Ns") {
            ns = read_float(lx);
        } else if (w == "Ni") {
            m.ior = read_float(lx);
        } else if (w == "d") {
            m.transparency = 1.0f - read_float(lx);
        } else if (w == "Tr") {
            m.transparency = read_float(lx);
        } else if (w == "map_Kd") {
            map(m.color_texture, true);
        } else if (w == "map_Pm") {
            map(m.metallic_texture, false);
        } else if (w == "map_Pr" || w == "map_roughness") {
            map(m.roughness_texture, false);
        } else if (w == "map_bump" || w == "norm" || w == "map_Kn") {
            map(m.normal_texture, false);
        }
        lx.skip_line();
    }
    build();
}

struct Face {
    int pos, norm, uv;
};

// ObjLoader.Parse (obj.cs:153-258) into the context's scene.
uint32_t parse_obj(sr_ctx* ctx, const std::string& path, const sr_transform& trans, const sr_material* defMat,
                   bool smooth) {
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) raise("OBJ: cannot read '" + path + "'");
    std::vector<float> positions, normals, uvs;  // 3, 3, 2 floats each
    std::vector<float> tp, tn, tu;              // per triangle: 9, 9, 6 floats
    std::vector<uint8_t> th;                    // per triangle: has vertex normals
    std::map<std::string, sr_material> materials;
    bool haveMat = false;
    std::string matName, objName;
    uint32_t objectCount = 0;

    auto build_object = [&] {
        const uint32_t n = (uint32_t)th.size();
        if (n == 0) return;
        bool anyN = false;
        for (uint8_t v : th) anyN = anyN || v;
        std::vector<float> tris((size_t)n * 36);
        sr::build_triangles(tp.data(), anyN ? tn.data() : nullptr, tu.data(), n, smooth && normals.empty(), tris.data(),
                            anyN ? th.data() : nullptr);
        sr_material mat;
        auto it = haveMat ? materials.find(matName) : materials.end();
        if (it != materials.end()) {
            mat = it->second;
        } else if (defMat) {
            mat = *defMat;
        } else {
            memset(&mat, 0, sizeof(mat));
            mat.albedo[0] = mat.albedo[1] = mat.albedo[2] = 1.0f;
            mat.roughness = 1.0f;
            mat.ior = 1.5f;
            mat.color_texture = mat.roughness_texture = mat.metallic_texture = mat.normal_texture = -1;
        }
        int32_t meshId = -1, objId = -1;
        if (sr_mesh_create(ctx, tris.data(), n, 8, &meshId) != SR_OK) raise(sr_last_error());
        if (sr_scene_add_object(ctx, &trans, &mat, SR_SHAPE_MESH, meshId, nullptr, &objId) != SR_OK)
            raise(sr_last_error());
        tp.clear();
        tn.clear();
        tu.clear();
        th.clear();
        ++objectCount;
    };

    Lexer lx(bytes.data(), bytes.size());
    std::string w;
    std::vector<Face> face;
    while (true) {
        const Tok k = lx.next(w, 512);
        if (k == Tok::End) break;
        if (k != Tok::Word) continue;
        if (w == "v") {
            for (int c = 0; c < 3; ++c) positions.push_back(read_float(lx));
            lx.skip_line();
        } else if (w == "vt") {
            const float u = read_float(lx), v = read_float(lx);
            uvs.push_back(u);
            uvs.push_back(1.0f - v);  // OBJ is bottom-left, textures top-left
            lx.skip_line();
        } else if (w == "vn") {
            for (int c = 0; c < 3; ++c) normals.push_back(read_float(lx));
            lx.skip_line();
        } else if (w == "f") {
            // ReadFace (obj.cs:444-468) then TriangulateFace (obj.cs:411-442).
            face.clear();
            const int np = (int)(positions.size() / 3), nn = (int)(normals.size() / 3), nu = (int)(uvs.size() / 2);
            while (lx.peek() == Tok::Word) {
                Face f{read_index(lx, np), -1, -1};
                if (lx.peek() == Tok::Slash) {
                    lx.next();
                    if (lx.peek() == Tok::Word) f.uv = read_index(lx, nu);
                    if (lx.peek() == Tok::Slash) {
                        lx.next();
                        if (lx.peek() == Tok::Word) f.norm = read_index(lx, nn);
                    }
                }
                face.push_back(f);
            }
            if (face.size() < 3) raise("OBJ: face has fewer than 3 vertices");
            auto P = [&](int i, int c) { return positions[(size_t)i * 3 + c]; };
            auto N = [&](int i, int c) { return normals[(size_t)i * 3 + c]; };
            const Face a = face[0];
            for (size_t i = 2; i < face.size(); ++i) {
                const Face b = face[i - 1], c = face[i];
                for (const Face* f : {&a, &b, &c})
                    for (int q = 0; q < 3; ++q) tp.push_back(P(f->pos, q));
                const float defU[3][2] = {{0.0f, 0.0f}, {1.0f, 0.0f}, {0.0f, 1.0f}};
                const Face* fs[3] = {&a, &b, &c};
                for (int j = 0; j < 3; ++j) {
                    if (fs[j]->uv >= 0) {
                        tu.push_back(uvs[(size_t)fs[j]->uv * 2]);
                        tu.push_back(uvs[(size_t)fs[j]->uv * 2 + 1]);
                    } else {
                        tu.push_back(defU[j][0]);
                        tu.push_back(defU[j][1]);
                    }
                }
                const bool hn = a.norm >= 0 && b.norm >= 0 && c.norm >= 0;
                for (int j = 0; j < 3; ++j)
                    for (int q = 0; q < 3; ++q) tn.push_back(hn ? N(fs[j]->norm, q) : 0.0f);
                th.push_back(hn ? 1 : 0);
            }
        } else if (w == "o" || w == "g") {
            build_object();
            if (lx.peek() == Tok::Word) lx.next(objName, 512);  // the object's name (descriptive only)
            lx.skip_line();
        } else if (w == "mtllib") {
            std::string f;
            lx.next(f, 512);
            const std::string mtlPath = combine(dir_of(path), f);
            if (!f.empty() && file_exists(mtlPath)) parse_mtl(ctx, mtlPath, materials);
            lx.skip_line();
        } else if (w == "usemtl") {
            build_object();
            lx.next(matName, 512);
            haveMat = true;
            lx.skip_line();
        } else {
            lx.skip_line();
        }
    }
    build_object();
    if (objectCount == 0) raise("OBJ: no objects found");
    return objectCount;
}

}  // namespace

extern "C" {

int sr_obj_load(sr_ctx* ctx, const char* path, const sr_transform* transform, const sr_material* material,
                int32_t smooth_normals, uint32_t* objects_added) {
    if (!ctx || !path) return sr_set_error(SR_ERR_INVALID, "obj: null argument");
    sr_transform t;
    if (transform) {
        t = *transform;
    } else {
        memset(&t, 0, sizeof(t));
        t.rot[3] = 1.0f;
        t.scale[0] = t.scale[1] = t.scale[2] = 1.0f;
    }
    try {
        const auto t0 = std::chrono::steady_clock::now();  // TimeObjLoad (obj.cs:147)
        const uint32_t n = parse_obj(ctx, path, t, material, smooth_normals != 0);
        sr::sr_count_time(ctx, SR_CT_TIME_OBJ_LOAD,
                          (int64_t)std::chrono::duration_cast<std::chrono::microseconds>(
                              std::chrono::steady_clock::now() - t0).count());
        if (objects_added) *objects_added = n;
        return SR_OK;
    } catch (const IngestError& e) {
        return sr_set_error(SR_ERR_INVALID, e.msg.c_str());
    } catch (const std::bad_alloc&) {
        return sr_set_error(SR_ERR_LIMIT, "obj: out of host memory");
    }
}

int sr_hdr_decode(const uint8_t* data, uint64_t size, float max_luminance, float* rgb, uint32_t* width,
                  uint32_t* height) {
    if (!data || !width || !height) return sr_set_error(SR_ERR_INVALID, "hdr: null argument");
    try {
        std::vector<uint8_t> rgbe;
        uint32_t w = 0, h = 0;
        decode_hdr(data, (size_t)size, rgbe, w, h);
        *width = w;
        *height = h;
        if (rgb) rgbe_to_float(rgbe, max_luminance, rgb);
        return SR_OK;
    } catch (const IngestError& e) {
        return sr_set_error(SR_ERR_INVALID, e.msg.c_str());
    } catch (const std::bad_alloc&) {
        return sr_set_error(SR_ERR_LIMIT, "hdr: out of host memory");
    }
}

int sr_hdr_load(const char* path, float max_luminance, float* rgb, uint32_t* width, uint32_t* height) {
    if (!path) return sr_set_error(SR_ERR_INVALID, "hdr: null path");
    std::vector<uint8_t> bytes;
    if (!read_file(path, bytes)) return sr_set_error(SR_ERR_INVALID, (std::string("hdr: cannot read ") + path).c_str());
    return sr_hdr_decode(bytes.data(), bytes.size(), max_luminance, rgb, width, height);
}

int sr_tga_decode(const uint8_t* data, uint64_t size, uint8_t* rgb, uint32_t* width, uint32_t* height) {
    if (!data || !width || !height) return sr_set_error(SR_ERR_INVALID, "tga: null argument");
    try {
        std::vector<uint8_t> px;
        uint32_t w = 0, h = 0;
        decode_tga(data, (size_t)size, px, w, h);
        *width = w;
        *height = h;
        if (rgb) memcpy(rgb, px.data(), px.size());
        return SR_OK;
    } catch (const IngestError& e) {
        return sr_set_error(SR_ERR_INVALID, e.msg.c_str());
    }
}

}  // extern "C"
