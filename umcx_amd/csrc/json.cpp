// json.cpp -- parser, accessors and printer for umcx::Json (see json.h).
#include "json.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace umcx {

namespace {

struct Parser {
    const char* p;
    const char* end;

    [[noreturn]] void fail(const char* what) const {
        throw JsonError(std::string("[json.exception.parse_error] ") + what + " at byte " +
                        std::to_string((long long)(p - (end - 0) + 0)));
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char* s) {
        size_t n = strlen(s);
        if ((size_t)(end - p) >= n && memcmp(p, s, n) == 0) { p += n; return true; }
        return false;
    }
    static void put_utf8(std::string& o, uint32_t cp) {
        if (cp < 0x80) o += (char)cp;
        else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) { o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F)); }
        else { o += (char)(0xF0 | (cp >> 18)); o += (char)(0x80 | ((cp >> 12) & 0x3F)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F)); }
    }
    uint32_t hex4() {
        if (end - p < 4) fail("truncated \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("invalid \\u escape");
        }
        return v;
    }
    std::string str() {
        if (p >= end || *p != '"') fail("expected string");
        ++p;
        std::string o;
        while (true) {
            if (p >= end) fail("unterminated string");
            char c = *p++;
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') { o += c; continue; }
            if (p >= end) fail("unterminated escape");
            char e = *p++;
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (!lit("\\u")) fail("unpaired surrogate");
                        uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("invalid escape");
            }
        }
        return o;
    }
    Json number() {
        const char* s = p;
        bool neg = false, is_float = false;
        if (*p == '-') { neg = true; ++p; }
        if (p >= end || !(*p >= '0' && *p <= '9')) fail("invalid number");
        while (p < end && *p >= '0' && *p <= '9') ++p;
        if (p < end && *p == '.') {
            is_float = true; ++p;
            if (p >= end || !(*p >= '0' && *p <= '9')) fail("invalid number");
            while (p < end && *p >= '0' && *p <= '9') ++p;
        }
        if (p < end && (*p == 'e' || *p == 'E')) {
            is_float = true; ++p;
            if (p < end && (*p == '+' || *p == '-')) ++p;
            if (p >= end || !(*p >= '0' && *p <= '9')) fail("invalid number");
            while (p < end && *p >= '0' && *p <= '9') ++p;
        }
        std::string tok(s, p);
        if (!is_float) {
            errno = 0;
            if (neg) {
                long long v = strtoll(tok.c_str(), nullptr, 10);
                if (errno != ERANGE) return Json((long long)v);
            } else {
                unsigned long long v = strtoull(tok.c_str(), nullptr, 10);
                if (errno != ERANGE) return Json((unsigned long long)v);
            }
        }
        return Json(strtod(tok.c_str(), nullptr));
    }
    Json value(int depth) {
        if (depth > 512) fail("nesting too deep");
        ws();
        if (p >= end) fail("unexpected end of input");
        char c = *p;
        if (c == '{') {
            ++p;
            Json o = Json::object();
            ws();
            if (p < end && *p == '}') { ++p; return o; }
            while (true) {
                ws();
                std::string k = str();
                ws();
                if (p >= end || *p != ':') fail("expected ':'");
                ++p;
                o[k] = value(depth + 1);
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return o; }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            Json a = Json::array();
            ws();
            if (p < end && *p == ']') { ++p; return a; }
            while (true) {
                a.push_back(value(depth + 1));
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return a; }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') return Json(str());
        if (lit("true")) return Json(true);
        if (lit("false")) return Json(false);
        if (lit("null")) return Json(nullptr);
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail("syntax error");
    }
};

void escape(std::string& out, const std::string& s) {
    out += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20) {
                    char buf[8];
                    snprintf(buf, sizeof buf, "\\u%04x", c);
                    out += buf;
                } else {
                    out += (char)c;
                }
        }
    }
    out += '"';
}

// shortest round-trip decimal, printed as fixed notation for exponents in [-5, 15) and with a
// trailing ".0" on integral values, otherwise d.ddde[+-]XX
void format_double(std::string& out, double d) {
    if (std::isnan(d) || std::isinf(d)) { out += "null"; return; }
    if (d == 0.0) { out += std::signbit(d) ? "-0.0" : "0.0"; return; }
    char buf[64];
    int prec = 1;
    for (; prec <= 17; prec++) {
        snprintf(buf, sizeof buf, "%.*e", prec - 1, d);
        if (strtod(buf, nullptr) == d) break;
    }
    // buf = [-]D.DDDDe[+-]XX
    std::string s(buf);
    bool neg = s[0] == '-';
    if (neg) s = s.substr(1);
    size_t epos = s.find('e');
    int exp10 = atoi(s.c_str() + epos + 1);
    std::string digits;
    for (size_t i = 0; i < epos; i++) if (s[i] != '.') digits += s[i];
    if (neg) out += '-';
    int n = (int)digits.size();
    if (exp10 >= -5 && exp10 < 15) {
        if (exp10 >= 0) {
            if (n <= exp10 + 1) {
                out += digits + std::string((size_t)(exp10 + 1 - n), '0') + ".0";
            } else {
                out += digits.substr(0, (size_t)exp10 + 1) + "." + digits.substr((size_t)exp10 + 1);
            }
        } else {
            out += "0." + std::string((size_t)(-exp10 - 1), '0') + digits;
        }
        return;
    }
    out += digits.substr(0, 1);
    if (n > 1) out += "." + digits.substr(1);
    snprintf(buf, sizeof buf, "e%c%02d", exp10 < 0 ? '-' : '+', exp10 < 0 ? -exp10 : exp10);
    out += buf;
}

}  // namespace

struct Json::Heap {
    std::string s;
    std::vector<std::string> keys;   // object keys (Object) -- parallel to vals
    std::vector<Json> vals;          // array elements (Array) or object values (Object)
};

const Json::Heap& Json::heap() const {
    static const Heap empty;
    return h_ ? *h_ : empty;
}
Json::Heap& Json::heap_mut() {
    if (!h_) h_ = new Heap;
    return *h_;
}
Json::Json(const char* s) : type_(Type::String) { heap_mut().s = s; }
Json::Json(std::string s) : type_(Type::String) { heap_mut().s = std::move(s); }
// the scalar union is copied as raw bytes, whichever member is active
Json::Json(const Json& o) : type_(o.type_), h_(o.h_ ? new Heap(*o.h_) : nullptr) { memcpy(&i_, &o.i_, sizeof i_); }
Json::Json(Json&& o) noexcept : type_(o.type_), h_(o.h_) {
    memcpy(&i_, &o.i_, sizeof i_);
    o.h_ = nullptr;
    o.type_ = Type::Null;
}
Json& Json::operator=(const Json& o) {
    if (this != &o) {
        Heap* h = o.h_ ? new Heap(*o.h_) : nullptr;   // copy first: o may live inside this
        delete h_;
        h_ = h;
        type_ = o.type_;
        memcpy(&i_, &o.i_, sizeof i_);
    }
    return *this;
}
Json& Json::operator=(Json&& o) noexcept {
    if (this != &o) {
        Heap* h = o.h_;
        const Type t = o.type_;
        int64_t bits;
        memcpy(&bits, &o.i_, sizeof bits);
        o.h_ = nullptr;
        o.type_ = Type::Null;
        delete h_;                                     // o may live inside this: taken out above
        h_ = h;
        type_ = t;
        memcpy(&i_, &bits, sizeof bits);
    }
    return *this;
}
const std::string& Json::key(size_t i) const { return heap().keys.at(i); }
const Json& Json::val(size_t i) const { return heap().vals.at(i); }
Json& Json::val(size_t i) { return heap_mut().vals.at(i); }
Json::~Json() { delete h_; }

Json Json::parse(const std::string& text) {
    Parser ps{text.data(), text.data() + text.size()};
    Json v = ps.value(0);
    ps.ws();
    if (ps.p != ps.end) ps.fail("unexpected trailing characters");
    return v;
}

double Json::as_double() const {
    switch (type_) {
        case Type::Int: return (double)i_;
        case Type::UInt: return (double)u_;
        case Type::Float: return d_;
        case Type::Bool: return b_ ? 1.0 : 0.0;
        default: throw JsonError("[json.exception.type_error.302] type must be number");
    }
}
int64_t Json::as_int() const {
    switch (type_) {
        case Type::Int: return i_;
        case Type::UInt: return (int64_t)u_;
        case Type::Float: return (int64_t)d_;
        case Type::Bool: return b_ ? 1 : 0;
        default: throw JsonError("[json.exception.type_error.302] type must be number");
    }
}
uint64_t Json::as_uint() const {
    switch (type_) {
        case Type::Int: return (uint64_t)i_;
        case Type::UInt: return u_;
        case Type::Float: return (uint64_t)d_;
        case Type::Bool: return b_ ? 1 : 0;
        default: throw JsonError("[json.exception.type_error.302] type must be number");
    }
}
const std::string& Json::as_string() const {
    if (type_ != Type::String) throw JsonError("[json.exception.type_error.302] type must be string");
    return heap().s;
}

size_t Json::size() const {
    if (type_ == Type::Array || type_ == Type::Object) return heap().vals.size();
    return type_ == Type::Null ? 0 : 1;
}
const Json* Json::find(const std::string& k) const {
    if (type_ != Type::Object) return nullptr;
    const Heap& h = heap();
    for (size_t i = 0; i < h.keys.size(); i++) if (h.keys[i] == k) return &h.vals[i];
    return nullptr;
}
Json* Json::find(const std::string& k) {
    if (type_ != Type::Object || !h_) return nullptr;
    for (size_t i = 0; i < h_->keys.size(); i++) if (h_->keys[i] == k) return &h_->vals[i];
    return nullptr;
}
bool Json::contains(const std::string& k) const { return find(k) != nullptr; }
const Json& Json::at(const std::string& k) const {
    const Json* v = find(k);
    if (!v) throw JsonError("[json.exception.out_of_range.403] key '" + k + "' not found");
    return *v;
}
const Json& Json::at(size_t i) const {
    if (type_ != Type::Array && type_ != Type::Object) throw JsonError("[json.exception.type_error.304] cannot use at() with a scalar");
    const Heap& h = heap();
    if (i >= h.vals.size()) throw JsonError("[json.exception.out_of_range.401] array index " + std::to_string(i) + " is out of range");
    return h.vals[i];
}
Json& Json::operator[](const std::string& k) {
    if (type_ == Type::Null) type_ = Type::Object;
    if (type_ != Type::Object) throw JsonError("[json.exception.type_error.305] cannot use operator[] with a string argument");
    Heap& h = heap_mut();
    for (size_t i = 0; i < h.keys.size(); i++) if (h.keys[i] == k) return h.vals[i];
    h.keys.push_back(k);
    h.vals.emplace_back();
    return h.vals.back();
}
Json& Json::operator[](size_t i) {
    if (type_ == Type::Null) type_ = Type::Array;
    if (type_ != Type::Array) throw JsonError("[json.exception.type_error.305] cannot use operator[] with a numeric argument");
    Heap& h = heap_mut();
    if (i >= h.vals.size()) h.vals.resize(i + 1);
    return h.vals[i];
}
void Json::push_back(Json v) {
    if (type_ == Type::Null) type_ = Type::Array;
    if (type_ != Type::Array) throw JsonError("[json.exception.type_error.308] cannot use push_back() with a non-array");
    heap_mut().vals.push_back(std::move(v));
}

double Json::value(const std::string& k, double dflt) const {
    const Json* v = find(k);
    return (v && !v->is_null()) ? v->as_double() : dflt;
}
int64_t Json::value_int(const std::string& k, int64_t dflt) const {
    const Json* v = find(k);
    return (v && !v->is_null()) ? v->as_int() : dflt;
}
std::string Json::value_str(const std::string& k, const std::string& dflt) const {
    const Json* v = find(k);
    return (v && !v->is_null()) ? v->as_string() : dflt;
}

void Json::update(const Json& other) {
    if (!other.is_object()) throw JsonError("[json.exception.type_error.312] cannot use update() with a non-object");
    if (type_ == Type::Null) type_ = Type::Object;
    if (type_ != Type::Object) throw JsonError("[json.exception.type_error.312] cannot use update() with a non-object");
    const Heap& oh = other.heap();
    for (size_t i = 0; i < oh.keys.size(); i++) {
        Json* mine = find(oh.keys[i]);
        if (mine && mine->is_object() && oh.vals[i].is_object()) {
            mine->update(oh.vals[i]);
        } else {
            (*this)[oh.keys[i]] = oh.vals[i];
        }
    }
}

void Json::set_pointer(const std::string& ptr, Json v) {
    if (ptr.empty()) { *this = std::move(v); return; }
    if (ptr[0] != '/') throw JsonError("[json.exception.parse_error.107] JSON pointer must be empty or begin with '/'");
    Json* cur = this;
    size_t pos = 1;
    while (true) {
        size_t nxt = ptr.find('/', pos);
        std::string tok = ptr.substr(pos, nxt == std::string::npos ? std::string::npos : nxt - pos);
        // unescape ~1 -> '/', ~0 -> '~'
        std::string key;
        for (size_t i = 0; i < tok.size(); i++) {
            if (tok[i] == '~' && i + 1 < tok.size()) { key += tok[i + 1] == '1' ? '/' : '~'; i++; }
            else key += tok[i];
        }
        Json& child = (cur->is_array()) ? (*cur)[(size_t)std::stoul(key)] : (*cur)[key];
        if (nxt == std::string::npos) { child = std::move(v); return; }
        cur = &child;
        pos = nxt + 1;
    }
}

void Json::dump_to(std::string& out, int indent, int level) const {
    char buf[32];
    switch (type_) {
        case Type::Null: out += "null"; return;
        case Type::Bool: out += b_ ? "true" : "false"; return;
        case Type::Int: snprintf(buf, sizeof buf, "%lld", (long long)i_); out += buf; return;
        case Type::UInt: snprintf(buf, sizeof buf, "%llu", (unsigned long long)u_); out += buf; return;
        case Type::Float: format_double(out, d_); return;
        case Type::String: escape(out, heap().s); return;
        case Type::Array:
        case Type::Object: {
            bool isobj = type_ == Type::Object;
            const Heap& h = heap();
            out += isobj ? '{' : '[';
            if (h.vals.empty()) { out += isobj ? '}' : ']'; return; }
            std::string pad = indent >= 0 ? std::string((size_t)indent * (level + 1), ' ') : std::string();
            for (size_t i = 0; i < h.vals.size(); i++) {
                if (i) out += ',';
                if (indent >= 0) { out += '\n'; out += pad; }
                if (isobj) {
                    escape(out, h.keys[i]);
                    out += indent >= 0 ? ": " : ":";
                }
                h.vals[i].dump_to(out, indent, level + 1);
            }
            if (indent >= 0) { out += '\n'; out += std::string((size_t)indent * level, ' '); }
            out += isobj ? '}' : ']';
            return;
        }
    }
}

std::string Json::dump(int indent) const {
    std::string out;
    dump_to(out, indent, 0);
    return out;
}

}  // namespace umcx
