// The ocean's four fields in the reference's save format (format 28), on the host: the CRC-32 and
// the chunk checksum, the directory of a grid, the four ocean records, and a reader and a rewriter
// of whole saves that go a chunk at a time over a save held in memory.
//
// The reader takes a frame, expands it (zlib's gzip for method 1), checks its checksum, then holds
// the part of every record prefix that falls inside the chunk to the bytes the directory fixes for
// it and copies out the part of the four ocean arrays that falls inside. The rewriter does the same
// and writes the chunk out again, with the ocean's elements written over, as a frame of the same
// index, so every other byte of the payload is carried and the chunks keep their places.
//
// Reference files are cited relative to cartography/src/commonMain/kotlin/com/cartogenesis/cartography/.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cgocean.h"
#include "cgsave.h"

namespace {

// The per-cell arrays of a save, in their order (WorldSections.SECTIONS, WorldSections.kt:144-233):
// the names and element types are the format's, so they are the reference's to the letter.
struct Spec {
    const char *name;
    int type;
};
constexpr int F32 = CG_SECTION_F32, I32 = CG_SECTION_I32, U8 = CG_SECTION_U8, UTF8 = CG_SECTION_UTF8;
constexpr Spec kSections[] = {
    {"terrain.normals.gradientX", F32}, {"terrain.normals.gradientY", F32},
    {"terrain.height", F32}, {"plates.plateId", I32},
    {"plates.boundaryDistance", F32}, {"plates.nearestBoundaryType", I32},
    {"plates.nearestBoundaryClass", I32}, {"plates.height", F32},
    {"plates.continentalShare", F32}, {"plates.seafloorAgeMyr", F32},
    {"plates.upliftRateMmPerYear", F32}, {"plates.crustAgeMyr", F32},
    {"plates.crustThicknessKm", F32}, {"plates.orogenicExcessKm", F32},
    {"plates.orogenyKind", I32}, {"plates.terraneId", I32},
    {"plates.riftAgeMyr", F32}, {"plates.rockErodibility", F32},
    {"plates.windowTectonicKm", F32}, {"plates.windowWearKm", F32},
    {"plates.windowOnsetMyrAgo", F32}, {"erosion.height", F32},
    {"sea.isLand", U8}, {"sea.relativeElevation", F32},
    {"ocean.velocityX", F32}, {"ocean.velocityY", F32},
    {"ocean.temperature", F32}, {"ocean.anomaly", F32},
    {"climate.temperature", F32}, {"climate.summerTemperature", F32},
    {"climate.winterTemperature", F32}, {"climate.julyTemperature", F32},
    {"climate.januaryTemperature", F32}, {"climate.precipitation", F32},
    {"climate.julyHalfPrecipitation", F32}, {"climate.januaryHalfPrecipitation", F32},
    {"climate.precipitationMm", F32}, {"climate.windDirection", I32},
    {"climate.julyHalfSeaIce", U8}, {"climate.januaryHalfSeaIce", U8},
    {"climate.windMeridional", F32}, {"climate.vegetationDensity", F32},
    {"climate.potentialEvapotranspirationMm", F32}, {"climate.openWaterEvaporationMm", F32},
    {"climate.windEastwardMps", F32}, {"climate.windSouthwardMps", F32},
    {"climate.permafrost", U8}, {"climate.biome", U8},
    {"rivers.filledElevation", F32}, {"rivers.flowAccumulation", F32},
    {"rivers.flowTarget", I32}, {"rivers.lakeId", I32},
    {"rivers.playa", U8}, {"nations.nationId", I32},
    {"nations.habitability", F32}, {"cultures.cultureId", I32},
};
constexpr int kArrays = (int)(sizeof(kSections) / sizeof(kSections[0]));
static_assert(kArrays + 1 == CG_SAVE_SECTIONS, "the directory is the lists and every array");
constexpr int kOceanFirst = 1 + 24;  // the directory entry of ocean.velocityX; the next three follow it
constexpr int kOceanFields = 4;
constexpr int kRecordPrefix = 4 + 4 + 4 + 8;       // WorldSections.RECORD_PREFIX_BYTES (WorldSections.kt:141)
constexpr int kPrefixBytes = 16;                   // WorldCodec.PREFIX_BYTES (WorldCodec.kt:343)
constexpr int kFrameBytes = 16;                    // PayloadWriter.FRAME_HEADER_BYTES (WorldSections.kt:643)
constexpr int64_t kLargestHeader = 16 << 20;       // WorldCodec.LARGEST_HEADER_BYTES (WorldCodec.kt:403)
constexpr int64_t kLargestLists = 64 << 20;        // WorldCodec.LARGEST_LISTS_BYTES (WorldCodec.kt:413)
constexpr int kUnfilledTail = 64;                  // UNFILLED_TAIL_BYTES (WorldSections.kt:894)
constexpr uint8_t kMagic[4] = {'C', 'G', 'W', 'D'};

int element_bytes(int type) { return type == U8 || type == UTF8 ? 1 : 4; }

const char *type_name(int type) {
    switch (type) {
    case F32: return "F32";
    case I32: return "I32";
    case U8: return "U8";
    default: return "UTF8";
    }
}

uint32_t get32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
void put32(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

// Crc32.of (PngWriter.kt:378-407) is the CRC-32 of zip and PNG, and a run continued from a checksum
// is zlib's crc32 of the same bytes from that checksum.
uint32_t crc_run(uint32_t continuing, const uint8_t *data, size_t len) {
    uLong crc = continuing;
    while (len > 0) {
        uInt step = len > (1u << 30) ? (1u << 30) : (uInt)len;
        crc = crc32(crc, data, step);
        data += step;
        len -= step;
    }
    return (uint32_t)crc;
}

uint32_t chunk_sum(uint32_t header_checksum, int index, const uint8_t *raw, size_t len) {
    uint8_t place[4];
    put32(place, (uint32_t)index);
    return crc_run(crc_run(header_checksum, place, 4), raw, len);
}

// A refusal: a SaveProblem and its detail, thrown inside this file and turned into a status at the
// C boundary.
struct Refusal {
    int code;
    std::string detail;
};

[[noreturn]] void refuse(int code, std::string detail) { throw Refusal{code, std::move(detail)}; }
[[noreturn]] void damaged(std::string detail) { refuse(CG_SAVE_DAMAGED, std::move(detail)); }
[[noreturn]] void incomplete(std::string detail) { refuse(CG_SAVE_INCOMPLETE, std::move(detail)); }

std::string str(int64_t v) { return std::to_string(v); }

std::vector<cg_section_info> directory_of(int cells, int lists_bytes) {
    std::vector<cg_section_info> out(CG_SAVE_SECTIONS);
    int64_t offset = 0;
    for (int i = 0; i < CG_SAVE_SECTIONS; ++i) {
        cg_section_info &e = out[i];
        std::memset(&e, 0, sizeof e);
        const char *name = i == 0 ? "lists" : kSections[i - 1].name;
        e.type = i == 0 ? UTF8 : kSections[i - 1].type;
        std::snprintf(e.name, sizeof e.name, "%s", name);
        e.count = i == 0 ? lists_bytes : cells;
        e.bytes = (int64_t)e.count * element_bytes(e.type);
        e.offset = offset;
        offset += kRecordPrefix + (int64_t)std::strlen(name) + e.bytes;
    }
    return out;
}

int64_t payload_of(const std::vector<cg_section_info> &dir) {
    int64_t total = 0;
    for (const auto &e : dir) total += kRecordPrefix + (int64_t)std::strlen(e.name) + e.bytes;
    return total;
}

// What PayloadWriter.record writes in front of a record's elements (WorldSections.kt:526-532).
std::vector<uint8_t> record_prefix(const cg_section_info &e) {
    size_t n = std::strlen(e.name);
    std::vector<uint8_t> p(kRecordPrefix + n);
    put32(p.data(), (uint32_t)n);
    std::memcpy(p.data() + 4, e.name, n);
    put32(p.data() + 4 + n, (uint32_t)e.type);
    put32(p.data() + 8 + n, (uint32_t)e.count);
    put32(p.data() + 12 + n, (uint32_t)(uint64_t)e.bytes);
    put32(p.data() + 16 + n, (uint32_t)((uint64_t)e.bytes >> 32));
    return p;
}

// ---- the header's JSON ---------------------------------------------------------------------------
// A reader of the one document the header is, enough to walk it: objects, arrays, strings (escapes
// decoded), numbers (kept as their text), true, false, null. Nesting is bounded so a hostile header
// cannot run the stack out.
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    std::string text;  // a string's value, a number's text, "true" or "false"
    std::vector<Json> items;
    std::vector<std::string> keys;  // an object's keys, in the order of items

    const Json *get(const char *key) const {
        if (kind != Object) return nullptr;
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] == key) return &items[i];
        return nullptr;
    }
};

class JsonReader {
  public:
    JsonReader(const char *p, const char *end) : p_(p), end_(end) {}

    Json document() {
        Json v = value(0);
        space();
        if (p_ != end_) fail("text after the document");
        return v;
    }

  private:
    const char *p_, *end_;

    [[noreturn]] void fail(const char *what) { damaged(std::string("its header does not parse: ") + what); }

    void space() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }

    bool take(char c) {
        space();
        if (p_ < end_ && *p_ == c) {
            ++p_;
            return true;
        }
        return false;
    }

    void expect(char c) {
        if (!take(c)) fail("a character out of place");
    }

    bool word(const char *w) {
        size_t n = std::strlen(w);
        if ((size_t)(end_ - p_) >= n && std::memcmp(p_, w, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }

    static void utf8(std::string &out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | cp >> 6);
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | cp >> 12);
            out += (char)(0x80 | (cp >> 6 & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | cp >> 18);
            out += (char)(0x80 | (cp >> 12 & 0x3F));
            out += (char)(0x80 | (cp >> 6 & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }

    uint32_t hex4() {
        if (end_ - p_ < 4) fail("a short \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("a bad \\u escape");
        }
        return v;
    }

    std::string string_body() {
        std::string out;
        for (;;) {
            if (p_ >= end_) fail("an unterminated string");
            char c = *p_++;
            if (c == '"') return out;
            if ((unsigned char)c < 0x20) fail("a control character in a string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= end_) fail("an unterminated escape");
            char e = *p_++;
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                uint32_t cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
                    p_ += 2;
                    uint32_t lo = hex4();
                    if (lo >= 0xDC00 && lo < 0xE000) cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    else fail("a lone surrogate");
                }
                utf8(out, cp);
                break;
            }
            default: fail("an unknown escape");
            }
        }
    }

    Json value(int depth) {
        if (depth > 64) fail("nesting too deep");
        space();
        if (p_ >= end_) fail("an early end");
        Json v;
        char c = *p_;
        if (c == '{') {
            ++p_;
            v.kind = Json::Object;
            if (take('}')) return v;
            do {
                expect('"');
                v.keys.push_back(string_body());
                expect(':');
                v.items.push_back(value(depth + 1));
            } while (take(','));
            expect('}');
        } else if (c == '[') {
            ++p_;
            v.kind = Json::Array;
            if (take(']')) return v;
            do {
                v.items.push_back(value(depth + 1));
            } while (take(','));
            expect(']');
        } else if (c == '"') {
            ++p_;
            v.kind = Json::String;
            v.text = string_body();
        } else if (word("true")) {
            v.kind = Json::Bool;
            v.text = "true";
        } else if (word("false")) {
            v.kind = Json::Bool;
            v.text = "false";
        } else if (word("null")) {
            v.kind = Json::Null;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            const char *s = p_;
            if (*p_ == '-') ++p_;
            while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' ||
                                 *p_ == '+' || *p_ == '-'))
                ++p_;
            v.kind = Json::Number;
            v.text.assign(s, p_);
        } else {
            fail("an unexpected character");
        }
        return v;
    }
};

// An integer field of the header: present, a number, integral and in range, or the header is damaged.
int64_t integer(const Json *v, const char *what, int64_t lo, int64_t hi) {
    if (!v || v->kind != Json::Number) damaged(std::string("its header has no ") + what);
    const std::string &t = v->text;
    size_t i = t[0] == '-' ? 1 : 0;
    if (i == t.size()) damaged(std::string("its header's ") + what + " is not a whole number");
    int64_t x = 0;
    for (; i < t.size(); ++i) {
        if (t[i] < '0' || t[i] > '9') damaged(std::string("its header's ") + what + " is not a whole number");
        if (x > (INT64_MAX - 9) / 10) damaged(std::string("its header's ") + what + " is out of range");
        x = x * 10 + (t[i] - '0');
    }
    if (t[0] == '-') x = -x;
    if (x < lo || x > hi) damaged(std::string("its header's ") + what + " is " + t);
    return x;
}

std::string describe(const cg_section_info &e) {
    return std::string("'") + e.name + "' (" + type_name(e.type) + ", " + str(e.count) + " elements, " + str(e.bytes) +
           " bytes at " + str(e.offset) + ")";
}

struct Header {
    uint32_t checksum = 0;
    size_t length = 0;  // the header's bytes
    int across = 0, down = 0;
    int64_t payload = 0;
    bool gzip = false, none = false;
    std::string compression;
    std::vector<cg_section_info> dir;
};

// WorldCodec.headerLengthFrom and headerFrom (WorldCodec.kt:575-670), for what the payload rests on:
// the prefix, the header's checksum, its version, its grid, and a directory equal, entry for entry,
// to the format's own for that grid. The document is otherwise carried as it is.
Header read_header(const uint8_t *file, size_t len) {
    if (len == 0) incomplete("the file is empty");
    size_t avail = len < (size_t)kPrefixBytes ? len : (size_t)kPrefixBytes;
    bool zeros = true;
    for (size_t i = 0; i < avail; ++i) zeros = zeros && file[i] == 0;
    if (zeros) incomplete("it holds nothing but zeros where a save begins");
    for (size_t i = 0; i < (avail < 4 ? avail : 4); ++i) {
        if (file[i] != kMagic[i]) {
            if (file[0] == '{')
                refuse(CG_SAVE_WRONG_VERSION, "it is a plain JSON save from before format 3, with no world in it; "
                                              "this build reads format 28");
            refuse(CG_SAVE_NOT_A_SAVE, "");
        }
    }
    if (len < (size_t)kPrefixBytes) incomplete("it ends after " + str((int64_t)len) + " bytes");
    int32_t version = (int32_t)get32(file + 4);
    if (version > CG_SAVE_FORMAT_VERSION)
        refuse(CG_SAVE_WRONG_VERSION, "a newer build wrote it, in format " + str(version) + "; this one reads 28");
    if (version < CG_SAVE_FORMAT_VERSION)
        refuse(CG_SAVE_WRONG_VERSION, "it is format " + str(version) + " and this build reads 28");
    int32_t hlen = (int32_t)get32(file + 8);
    if (hlen <= 0) damaged("its header is " + str(hlen) + " bytes long");
    if (hlen > kLargestHeader) refuse(CG_SAVE_TOO_LARGE, "its header claims " + str(hlen) + " bytes");
    if (len - kPrefixBytes < (size_t)hlen) incomplete("it ends inside its header");

    Header h;
    h.length = (size_t)hlen;
    h.checksum = get32(file + 12);
    const uint8_t *text = file + kPrefixBytes;
    if (crc_run(0, text, h.length) != h.checksum) damaged("its header fails its checksum");
    Json doc = JsonReader((const char *)text, (const char *)text + h.length).document();
    if (doc.kind != Json::Object) damaged("its header does not parse: not an object");
    int64_t fv = integer(doc.get("formatVersion"), "formatVersion", INT32_MIN, INT32_MAX);
    if (fv != CG_SAVE_FORMAT_VERSION) damaged("its header says format " + str(fv) + " where its prefix says 28");
    const Json *document = doc.get("document");
    const Json *config = document ? document->get("config") : nullptr;
    if (!config) damaged("its header has no settings");
    int64_t width = integer(config->get("width"), "width", INT32_MIN, INT32_MAX);
    int64_t height = integer(config->get("height"), "height", INT32_MIN, INT32_MAX);
    if (width <= 0 || height <= 0) damaged("its grid is " + str(width) + " by " + str(height));
    if (width * height > CG_SAVE_LARGEST_GRID_CELLS)
        refuse(CG_SAVE_TOO_LARGE, "its grid is " + str(width) + " by " + str(height));
    h.across = (int)width;
    h.down = (int)height;

    const Json *comp = doc.get("compression");
    if (!comp || comp->kind != Json::String) damaged("its header names no compression");
    h.compression = comp->text;
    h.gzip = h.compression == "gzip";
    h.none = h.compression == "none";

    const Json *sections = doc.get("sections");
    if (!sections || sections->kind != Json::Array) damaged("its header has no directory");
    if (sections->items.empty()) damaged("its directory is empty");
    std::vector<cg_section_info> found;
    for (const Json &s : sections->items) {
        cg_section_info e;
        std::memset(&e, 0, sizeof e);
        const Json *name = s.get("name");
        const Json *type = s.get("type");
        if (!name || name->kind != Json::String || !type || type->kind != Json::String)
            damaged("its header does not parse: a directory entry without a name or a type");
        std::snprintf(e.name, sizeof e.name, "%s", name->text.c_str());
        e.type = type->text == "F32" ? F32 : type->text == "I32" ? I32 : type->text == "U8" ? U8
                 : type->text == "UTF8" ? UTF8 : 0;
        e.count = (int)integer(s.get("count"), "count", INT32_MIN, INT32_MAX);
        e.bytes = integer(s.get("bytes"), "bytes", INT64_MIN + 1, INT64_MAX);
        e.offset = integer(s.get("offset"), "offset", INT64_MIN + 1, INT64_MAX);
        if (name->text.size() >= sizeof e.name || e.type == 0) {
            damaged("entry " + str((int64_t)found.size()) + " of its directory is '" + name->text + "' (" + type->text +
                    ") where this build's format has none such");
        }
        found.push_back(e);
    }
    const cg_section_info &lists = found[0];
    if (std::strcmp(lists.name, "lists") == 0 && lists.count > kLargestLists)
        refuse(CG_SAVE_TOO_LARGE, "its lists claim " + str(lists.count) + " bytes");
    h.dir = directory_of((int)(width * height), lists.count > 0 ? lists.count : 0);
    if (found.size() != h.dir.size())
        damaged("its directory lists " + str((int64_t)found.size()) + " sections where this build's format has " +
                str((int64_t)h.dir.size()));
    for (size_t i = 0; i < found.size(); ++i) {
        const cg_section_info &a = found[i], &b = h.dir[i];
        if (std::strcmp(a.name, b.name) != 0 || a.type != b.type || a.count != b.count || a.bytes != b.bytes ||
            a.offset != b.offset)
            damaged("entry " + str((int64_t)i) + " of its directory is " + describe(a) + " where this build's format has " +
                    describe(b));
    }
    h.payload = payload_of(h.dir);
    int64_t promised = integer(doc.get("payloadBytes"), "payloadBytes", INT64_MIN + 1, INT64_MAX);
    if (promised != h.payload)
        damaged("its header promises " + str(promised) + " payload bytes where its directory adds up to " +
                str(h.payload));
    return h;
}

bool ends_unfilled(const uint8_t *b, size_t n) {
    if (n < (size_t)kUnfilledTail) return false;
    for (size_t i = n - kUnfilledTail; i < n; ++i)
        if (b[i] != 0) return false;
    return true;
}

// The expanded payload coming back a chunk at a time (PayloadReader.nextFrame, WorldSections.kt:799-842).
class Frames {
  public:
    Frames(const uint8_t *file, size_t len, const Header &h)
        : file_(file), len_(len), h_(h), pos_(kPrefixBytes + h.length), chunk_(CG_SAVE_CHUNK_BYTES + 1) {}
    ~Frames() {
        if (inflating_) inflateEnd(&zs_);
    }

    // The next chunk into chunk(), or false once the directory's bytes are all delivered.
    bool next() {
        if (delivered_ == h_.payload) return false;
        const std::string where = "after " + str(delivered_) + " of " + str(h_.payload) + " bytes";
        if (len_ - pos_ < (size_t)kFrameBytes) incomplete("it ends partway through its world, " + where);
        const uint8_t *fh = file_ + pos_;
        bool zeros = true;
        for (int i = 0; i < kFrameBytes; ++i) zeros = zeros && fh[i] == 0;
        if (zeros) incomplete("it holds zeros where its world should continue, " + where);
        int32_t raw = (int32_t)get32(fh), stored = (int32_t)get32(fh + 4), method = (int32_t)get32(fh + 12);
        uint32_t checksum = get32(fh + 8);
        const std::string f = "chunk " + str(frames_);
        if (raw < 1 || raw > CG_SAVE_CHUNK_BYTES) damaged(f + " claims " + str(raw) + " bytes");
        if (delivered_ + raw > h_.payload) damaged(f + " runs past the " + str(h_.payload) + " bytes the directory promises");
        pos_ += kFrameBytes;
        if (method == 0) {
            if (stored != raw) damaged(f + " is stored raw but claims " + str(stored) + " bytes for " + str(raw));
            body(stored, where);
            std::memcpy(chunk_.data(), file_ + pos_, (size_t)raw);
        } else if (method == 1) {
            if (stored < 1 || stored >= raw) damaged(f + " claims " + str(stored) + " compressed bytes for " + str(raw));
            body(stored, where);
            expand(file_ + pos_, (size_t)stored, (size_t)raw, f, where);
        } else {
            damaged(f + " names an unknown method " + str(method));
        }
        pos_ += (size_t)stored;
        if (chunk_sum(h_.checksum, frames_, chunk_.data(), (size_t)raw) != checksum) {
            if (ends_unfilled(chunk_.data(), (size_t)raw))
                incomplete("it holds zeros where its world should continue, in " + f + " " + where);
            damaged(f + " fails its checksum");
        }
        start_ = delivered_;
        size_ = (size_t)raw;
        delivered_ += raw;
        ++frames_;
        return true;
    }

    // PayloadReader.finish (WorldSections.kt:788-797): the closing frame and nothing after it.
    void finish() {
        if (len_ - pos_ < (size_t)kFrameBytes) incomplete("it ends where the frame that closes the world should be");
        for (int i = 0; i < kFrameBytes; ++i)
            if (file_[pos_ + i] != 0) damaged("the world runs on past where its directory ends");
        if (len_ - pos_ > (size_t)kFrameBytes) damaged("there are bytes after the end of the world");
    }

    uint8_t *chunk() { return chunk_.data(); }
    size_t size() const { return size_; }
    int64_t start() const { return start_; }  // the chunk's first byte in the expanded payload
    int index() const { return frames_ - 1; }

  private:
    const uint8_t *file_;
    size_t len_;
    const Header &h_;
    size_t pos_;
    int frames_ = 0;
    int64_t delivered_ = 0, start_ = 0;
    size_t size_ = 0;
    std::vector<uint8_t> chunk_;
    z_stream zs_{};
    bool inflating_ = false;

    void body(int32_t stored, const std::string &where) {
        if (len_ - pos_ < (size_t)stored) incomplete("it ends partway through chunk " + str(frames_) + ", " + where);
    }

    // A gzip chunk expanded to exactly its raw length (WorldSections.kt:864-887); members one after
    // another expand as one run.
    void expand(const uint8_t *in, size_t n, size_t raw, const std::string &f, const std::string &where) {
        if (h_.none) damaged(f + " is compressed in a file whose header says it stores its chunks raw");
        if (!h_.gzip)
            refuse(CG_SAVE_CANNOT_EXPAND,
                   "its chunks are compressed with '" + h_.compression + "' and this platform expands 'gzip'");
        if (!inflating_) {
            if (inflateInit2(&zs_, 15 + 16) != Z_OK) refuse(CG_SAVE_CANNOT_EXPAND, "this platform cannot expand 'gzip'");
            inflating_ = true;
        }
        size_t out = 0;
        const uint8_t *p = in;
        size_t left = n;
        for (;;) {
            inflateReset(&zs_);
            zs_.next_in = (Bytef *)p;
            zs_.avail_in = (uInt)left;
            zs_.next_out = chunk_.data() + out;
            zs_.avail_out = (uInt)(raw + 1 - out);
            int rc = inflate(&zs_, Z_FINISH);
            out = (size_t)(zs_.next_out - chunk_.data());
            if (out > raw) damaged(f + " expands to more than where its frame says " + str((int64_t)raw) + " bytes");
            if (rc != Z_STREAM_END) {
                if (ends_unfilled(in, n))
                    incomplete("it holds zeros where its world should continue, in " + f + " " + where);
                damaged(f + " does not expand: " + (zs_.msg ? zs_.msg : "the stream ends early"));
            }
            p = zs_.next_in;
            left = zs_.avail_in;
            if (left < 2 || p[0] != 0x1f || p[1] != 0x8b) break;
        }
        if (out != raw)
            damaged(f + " expands to " + str((int64_t)out) + " where its frame says " + str((int64_t)raw) + " bytes");
    }
};

// Every record prefix's bytes and every ocean array's element range, in payload coordinates.
struct Layout {
    std::vector<std::vector<uint8_t>> prefixes;
    std::vector<int64_t> at;
    int64_t ocean_at[kOceanFields];  // the first element byte of each ocean array
    int64_t ocean_bytes;

    explicit Layout(const Header &h) {
        for (const auto &e : h.dir) {
            prefixes.push_back(record_prefix(e));
            at.push_back(e.offset);
        }
        for (int k = 0; k < kOceanFields; ++k) {
            const cg_section_info &e = h.dir[kOceanFirst + k];
            ocean_at[k] = e.offset + kRecordPrefix + (int64_t)std::strlen(e.name);
        }
        ocean_bytes = h.dir[kOceanFirst].bytes;
    }

    // PayloadReader.expectRecord (WorldSections.kt:676-690) for what of every prefix lies in the chunk.
    void check(const uint8_t *chunk, int64_t start, size_t size, const Header &h) const {
        int64_t end = start + (int64_t)size;
        for (size_t i = 0; i < prefixes.size(); ++i) {
            int64_t a = at[i], b = at[i] + (int64_t)prefixes[i].size();
            if (b <= start || a >= end) continue;
            int64_t lo = a > start ? a : start, hi = b < end ? b : end;
            if (std::memcmp(chunk + (lo - start), prefixes[i].data() + (lo - a), (size_t)(hi - lo)) != 0)
                damaged("record " + str((int64_t)i) + " does not begin as its directory entry " + describe(h.dir[i]) +
                        " says");
        }
    }

    // Calls fn(field, chunk offset, field byte offset, bytes) for what of each ocean array lies in the chunk.
    template <class Fn> void ocean(int64_t start, size_t size, Fn fn) const {
        int64_t end = start + (int64_t)size;
        for (int k = 0; k < kOceanFields; ++k) {
            int64_t a = ocean_at[k], b = a + ocean_bytes;
            if (b <= start || a >= end) continue;
            int64_t lo = a > start ? a : start, hi = b < end ? b : end;
            fn(k, (size_t)(lo - start), (size_t)(lo - a), (size_t)(hi - lo));
        }
    }
};

std::string float_text(float v) {
    if (std::isnan(v)) return "NaN";
    return v > 0 ? "Infinity" : "-Infinity";
}

// The first value of an ocean array that is not finite (WorldSections.kt:744-747), or none.
// An array the caller hands in is a bad argument; one read from a save is damage.
void check_finite(const float *const fields[kOceanFields], int64_t cells, bool input) {
    for (int k = 0; k < kOceanFields; ++k) {
        for (int64_t c = 0; c < cells; ++c) {
            if (std::isfinite(fields[k][c])) continue;
            std::string name = kSections[kOceanFirst - 1 + k].name;
            if (input) refuse(CG_ERR_ARG, "the new '" + name + "' holds " + float_text(fields[k][c]) + " at cell " + str(c));
            damaged("'" + name + "' holds " + float_text(fields[k][c]) + " at cell " + str(c));
        }
    }
}

void set_reason(char *reason, size_t n, const std::string &text) {
    if (reason && n > 0) std::snprintf(reason, n, "%s", text.c_str());
}

template <class Fn> int guarded(char *reason, size_t reason_len, Fn fn) {
    try {
        fn();
        set_reason(reason, reason_len, "");
        return CG_OK;
    } catch (const Refusal &r) {
        set_reason(reason, reason_len, r.detail);
        return r.code;
    } catch (const std::bad_alloc &) {
        set_reason(reason, reason_len, "out of host memory");
        return CG_ERR_NOMEM;
    }
}

// One chunk written out as a frame (PayloadWriter.flush, WorldSections.kt:627-640).
class FrameWriter {
  public:
    FrameWriter(uint8_t *out, size_t cap, bool compress, uint32_t header_checksum)
        : out_(out), cap_(cap), compress_(compress), checksum_(header_checksum), packed_(CG_SAVE_CHUNK_BYTES) {}
    ~FrameWriter() {
        if (deflating_) deflateEnd(&zs_);
    }

    void put(const uint8_t *b, size_t n) {
        if (cap_ - pos_ < n) refuse(CG_ERR_ARG, "the output holds " + str((int64_t)cap_) + " bytes, too few for the save");
        std::memcpy(out_ + pos_, b, n);
        pos_ += n;
    }

    void frame(int index, const uint8_t *raw, size_t n) {
        size_t packed = compress_ ? deflate_chunk(raw, n) : 0;
        uint8_t fh[kFrameBytes];
        put32(fh, (uint32_t)n);
        put32(fh + 4, (uint32_t)(packed ? packed : n));
        put32(fh + 8, chunk_sum(checksum_, index, raw, n));
        put32(fh + 12, packed ? 1u : 0u);
        put(fh, kFrameBytes);
        put(packed ? packed_.data() : raw, packed ? packed : n);
    }

    void close() {
        uint8_t zeros[kFrameBytes] = {0};
        put(zeros, kFrameBytes);
    }

    size_t written() const { return pos_; }

  private:
    uint8_t *out_;
    size_t cap_, pos_ = 0;
    bool compress_;
    uint32_t checksum_;
    std::vector<uint8_t> packed_;
    z_stream zs_{};
    bool deflating_ = false;

    // The chunk as one gzip member when that is shorter than the chunk, else 0.
    size_t deflate_chunk(const uint8_t *raw, size_t n) {
        if (!deflating_) {
            if (deflateInit2(&zs_, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK)
                refuse(CG_SAVE_CANNOT_EXPAND, "this platform cannot compress with 'gzip'");
            deflating_ = true;
        }
        deflateReset(&zs_);
        // The member's header as the JVM's gzip writes it: no name, no time, the system "unknown".
        gz_header gh{};
        gh.os = 255;
        deflateSetHeader(&zs_, &gh);
        zs_.next_in = (Bytef *)raw;
        zs_.avail_in = (uInt)n;
        zs_.next_out = packed_.data();
        zs_.avail_out = (uInt)(n - 1);
        int rc = deflate(&zs_, Z_FINISH);
        return rc == Z_STREAM_END ? (size_t)(zs_.next_out - packed_.data()) : 0;
    }
};

// The whole walk: every chunk checked, the ocean copied out of it (read) or written over in it and
// the chunk framed again (replace).
void walk(const uint8_t *file, size_t len, const Header &h, float *const read_into[kOceanFields],
          const float *const write_from[kOceanFields], FrameWriter *writer) {
    Layout lay(h);
    int64_t cells = (int64_t)h.across * h.down;
    if (write_from) check_finite(write_from, cells, true);
    if (writer) {
        writer->put(file, kPrefixBytes + h.length);
    }
    Frames frames(file, len, h);
    while (frames.next()) {
        uint8_t *chunk = frames.chunk();
        lay.check(chunk, frames.start(), frames.size(), h);
        lay.ocean(frames.start(), frames.size(), [&](int k, size_t at, size_t field_at, size_t n) {
            if (read_into) std::memcpy((uint8_t *)read_into[k] + field_at, chunk + at, n);
            if (write_from) std::memcpy(chunk + at, (const uint8_t *)write_from[k] + field_at, n);
        });
        if (writer) writer->frame(frames.index(), chunk, frames.size());
    }
    frames.finish();
    if (read_into) {
        const float *fields[kOceanFields] = {read_into[0], read_into[1], read_into[2], read_into[3]};
        check_finite(fields, cells, false);
    }
    if (writer) writer->close();
}

}  // namespace

extern "C" {

uint32_t cg_save_crc32(const uint8_t *data, size_t len, uint32_t continuing) {
    if (!data && len) return 0;
    return crc_run(continuing, data, len);
}

uint32_t cg_save_chunk_checksum(uint32_t header_checksum, int index, const uint8_t *raw, size_t len) {
    if (!raw && len) return 0;
    return chunk_sum(header_checksum, index, raw, len);
}

int cg_save_directory(int cells, int lists_bytes, cg_section_info *out, int cap, int64_t *payload_bytes) {
    if (cells <= 0 || (int64_t)cells > CG_SAVE_LARGEST_GRID_CELLS || lists_bytes < 0 || cap < 0 || (cap > 0 && !out))
        return CG_ERR_ARG;
    std::vector<cg_section_info> dir = directory_of(cells, lists_bytes);
    for (int i = 0; i < cap && i < CG_SAVE_SECTIONS; ++i) out[i] = dir[i];
    if (payload_bytes) *payload_bytes = payload_of(dir);
    return CG_SAVE_SECTIONS;
}

int cg_ocean_records_write(int cells, const float *velocity_x, const float *velocity_y, const float *temperature,
                           const float *anomaly, uint8_t *out, size_t cap, size_t *written) {
    const float *fields[kOceanFields] = {velocity_x, velocity_y, temperature, anomaly};
    if (cells <= 0 || (int64_t)cells > CG_SAVE_LARGEST_GRID_CELLS || !out) return CG_ERR_ARG;
    for (const float *f : fields)
        if (!f) return CG_ERR_ARG;
    std::vector<cg_section_info> dir = directory_of(cells, 0);
    size_t need = 0;
    for (int k = 0; k < kOceanFields; ++k) need += kRecordPrefix + std::strlen(dir[kOceanFirst + k].name) + 4 * (size_t)cells;
    if (written) *written = need;
    if (cap < need) return CG_ERR_ARG;
    for (int k = 0; k < kOceanFields; ++k)
        for (int c = 0; c < cells; ++c)
            if (!std::isfinite(fields[k][c])) return CG_ERR_ARG;
    size_t pos = 0;
    for (int k = 0; k < kOceanFields; ++k) {
        std::vector<uint8_t> p = record_prefix(dir[kOceanFirst + k]);
        std::memcpy(out + pos, p.data(), p.size());
        pos += p.size();
        for (int c = 0; c < cells; ++c) {
            uint32_t bits;
            std::memcpy(&bits, fields[k] + c, 4);
            put32(out + pos, bits);
            pos += 4;
        }
    }
    return CG_OK;
}

int cg_ocean_records_read(const uint8_t *records, size_t len, int cells, float *velocity_x, float *velocity_y,
                          float *temperature, float *anomaly, char *reason, size_t reason_len) {
    float *fields[kOceanFields] = {velocity_x, velocity_y, temperature, anomaly};
    if (cells <= 0 || (int64_t)cells > CG_SAVE_LARGEST_GRID_CELLS || (!records && len)) return CG_ERR_ARG;
    for (float *f : fields)
        if (!f) return CG_ERR_ARG;
    return guarded(reason, reason_len, [&] {
        std::vector<cg_section_info> dir = directory_of(cells, 0);
        size_t pos = 0;
        for (int k = 0; k < kOceanFields; ++k) {
            const cg_section_info &e = dir[kOceanFirst + k];
            std::vector<uint8_t> p = record_prefix(e);
            if (len - pos < p.size() + 4 * (size_t)cells)
                incomplete("the records end inside '" + std::string(e.name) + "'");
            if (std::memcmp(records + pos, p.data(), p.size()) != 0)
                damaged("record " + str(kOceanFirst + k) + " does not begin as its directory entry " + describe(e) + " says");
            pos += p.size();
            for (int c = 0; c < cells; ++c) {
                uint32_t bits = get32(records + pos);
                std::memcpy(fields[k] + c, &bits, 4);
                pos += 4;
            }
        }
        const float *view[kOceanFields] = {velocity_x, velocity_y, temperature, anomaly};
        check_finite(view, cells, false);
        if (pos != len) damaged("the records run " + str((int64_t)(len - pos)) + " bytes past the last ocean section");
    });
}

int cg_save_inspect(const uint8_t *file, size_t len, int *across, int *down, int64_t *payload_bytes,
                    int64_t *largest_rewrite, char *reason, size_t reason_len) {
    if (!file && len) return CG_ERR_ARG;
    return guarded(reason, reason_len, [&] {
        Header h = read_header(file, len);
        // The rewrite keeps the input's chunks, so it has as many frames: count them from their headers.
        int64_t frames = 0, delivered = 0;
        size_t pos = kPrefixBytes + h.length;
        while (delivered < h.payload && len - pos >= (size_t)kFrameBytes) {
            int32_t raw = (int32_t)get32(file + pos), stored = (int32_t)get32(file + pos + 4);
            if (raw < 1 || stored < 1) break;
            delivered += raw;
            ++frames;
            pos += kFrameBytes + (size_t)stored;
            if (pos > len) break;
        }
        if (across) *across = h.across;
        if (down) *down = h.down;
        if (payload_bytes) *payload_bytes = h.payload;
        if (largest_rewrite) *largest_rewrite = kPrefixBytes + (int64_t)h.length + h.payload + kFrameBytes * (frames + 1);
    });
}

int cg_save_read_ocean(const uint8_t *file, size_t len, float *velocity_x, float *velocity_y, float *temperature,
                       float *anomaly, char *reason, size_t reason_len) {
    float *fields[kOceanFields] = {velocity_x, velocity_y, temperature, anomaly};
    if (!file && len) return CG_ERR_ARG;
    for (float *f : fields)
        if (!f) return CG_ERR_ARG;
    return guarded(reason, reason_len, [&] { walk(file, len, read_header(file, len), fields, nullptr, nullptr); });
}

int cg_save_replace_ocean(const uint8_t *file, size_t len, const float *velocity_x, const float *velocity_y,
                          const float *temperature, const float *anomaly, int store, uint8_t *out, size_t cap,
                          size_t *written, char *reason, size_t reason_len) {
    const float *fields[kOceanFields] = {velocity_x, velocity_y, temperature, anomaly};
    if ((!file && len) || !out) return CG_ERR_ARG;
    for (const float *f : fields)
        if (!f) return CG_ERR_ARG;
    if (written) *written = 0;
    return guarded(reason, reason_len, [&] {
        Header h = read_header(file, len);
        FrameWriter writer(out, cap, h.gzip && !store, h.checksum);
        walk(file, len, h, nullptr, fields, &writer);
        if (written) *written = writer.written();
    });
}

}  // extern "C"
