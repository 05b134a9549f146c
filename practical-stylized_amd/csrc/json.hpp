// Minimal JSON reader shared by the config parser and the scene importer: a recursive-descent
// parser into a tree of values (objects keep their key order).
#pragma once
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace nlpf {

struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
};

struct JParser {
    const char* s;
    std::string err;
    void ws() {
        while (*s == ' ' || *s == '\n' || *s == '\t' || *s == '\r') ++s;
    }
    bool fail(const std::string& m) {
        if (err.empty()) err = m;
        return false;
    }
    bool parse_string(std::string& out) {
        if (*s != '"') return fail("expected string");
        ++s;
        while (*s && *s != '"') {
            if (*s == '\\') {
                ++s;
                switch (*s) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        // keep ASCII code points, replace others
                        unsigned v = 0;
                        for (int i = 0; i < 4; ++i) {
                            ++s;
                            char c = *s;
                            if (!c) return fail("bad \\u escape");
                            v = v * 16 + (unsigned)(isdigit((unsigned char)c) ? c - '0' : (tolower(c) - 'a' + 10));
                        }
                        out += v < 128 ? (char)v : '?';
                        break;
                    }
                    case 0: return fail("unterminated escape");
                    default: out += *s;
                }
                ++s;
            } else {
                out += *s++;
            }
        }
        if (*s != '"') return fail("unterminated string");
        ++s;
        return true;
    }
    int depth = 0;  // nesting of the value being parsed
    bool parse(JVal& v) {
        if (depth >= 256) return fail("JSON nested deeper than 256 levels");
        ++depth;
        const bool ok = parse_value(v);
        --depth;
        return ok;
    }
    bool parse_value(JVal& v) {
        ws();
        if (*s == '{') {
            ++s;
            v.kind = JVal::Obj;
            ws();
            if (*s == '}') {
                ++s;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!parse_string(k)) return false;
                ws();
                if (*s != ':') return fail("expected ':'");
                ++s;
                JVal c;
                if (!parse(c)) return false;
                v.obj.emplace_back(k, std::move(c));
                ws();
                if (*s == ',') {
                    ++s;
                    continue;
                }
                if (*s == '}') {
                    ++s;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (*s == '[') {
            ++s;
            v.kind = JVal::Arr;
            ws();
            if (*s == ']') {
                ++s;
                return true;
            }
            for (;;) {
                JVal c;
                if (!parse(c)) return false;
                v.arr.push_back(std::move(c));
                ws();
                if (*s == ',') {
                    ++s;
                    continue;
                }
                if (*s == ']') {
                    ++s;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (*s == '"') {
            v.kind = JVal::Str;
            return parse_string(v.str);
        }
        if (!strncmp(s, "true", 4)) {
            s += 4;
            v.kind = JVal::Bool;
            v.b = true;
            return true;
        }
        if (!strncmp(s, "false", 5)) {
            s += 5;
            v.kind = JVal::Bool;
            v.b = false;
            return true;
        }
        if (!strncmp(s, "null", 4)) {
            s += 4;
            v.kind = JVal::Null;
            return true;
        }
        char* end = nullptr;
        double d = strtod(s, &end);
        if (end == s) return fail(std::string("unexpected character '") + (*s ? *s : '0') + "'");
        s = end;
        v.kind = JVal::Num;
        v.num = d;
        return true;
    }
};

// member of an object, nullptr when absent (or when v is not an object)
inline const JVal* jget(const JVal& v, const char* key) {
    if (v.kind != JVal::Obj) return nullptr;
    for (const auto& kv : v.obj)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

}  // namespace nlpf
