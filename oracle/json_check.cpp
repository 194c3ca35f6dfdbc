/*
 * json_check.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * Differential check of the scene files' JSON reader (the library's
 * pico_amd/csrc/json_reader.h) against jsoncpp, the reader of the reference's
 * scene loader (Source/Core/Scene.cpp:43-47: std::istream >> Json::Value with
 * the default CharReaderBuilder settings), compiled from the reference tree
 * by oracle/Makefile into oracle/_ref/libjson_check.so.
 *
 * Both parse the same text and print their value tree in one canonical form:
 * one line per node, "path kind value", with strings and member names as
 * hexadecimal bytes and numbers as the exact double and float (%a) that
 * asDouble / asFloat give.  A parse failure prints nothing and returns 0.
 */
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>

#include "json/json.h"

#include "../pico_amd/csrc/json_reader.h"

namespace {

std::string hex(const std::string& s) {
    static const char* d = "0123456789abcdef";
    std::string o;
    for (unsigned char c : s) {
        o += d[c >> 4];
        o += d[c & 15];
    }
    return o;
}

std::string num(double x, float f) {
    char b[96];
    std::snprintf(b, sizeof b, "%a %a", x, (double)f);
    return b;
}

void flat_ref(const Json::Value& v, const std::string& path, std::string& out) {
    switch (v.type()) {
        case Json::nullValue: out += path + " null\n"; break;
        case Json::booleanValue: out += path + (v.asBool() ? " true\n" : " false\n"); break;
        case Json::intValue:
        case Json::uintValue:
        case Json::realValue: out += path + " num " + num(v.asDouble(), v.asFloat()) + "\n"; break;
        case Json::stringValue: out += path + " str " + hex(v.asString()) + "\n"; break;
        case Json::arrayValue:
            out += path + " arr " + std::to_string(v.size()) + "\n";
            for (Json::ArrayIndex i = 0; i < v.size(); ++i) flat_ref(v[i], path + "[" + std::to_string(i) + "]", out);
            break;
        case Json::objectValue:
            out += path + " obj " + std::to_string(v.size()) + "\n";
            for (auto it = v.begin(); it != v.end(); ++it) flat_ref(*it, path + "." + hex(it.name()), out);
            break;
    }
}

void flat_pico(const pico_json::Value& v, const std::string& path, std::string& out) {
    using V = pico_json::Value;
    switch (v.kind) {
        case V::Null: out += path + " null\n"; break;
        case V::Bool: out += path + (v.b ? " true\n" : " false\n"); break;
        case V::Num: out += path + " num " + num(v.as_double(), v.as_float()) + "\n"; break;
        case V::Str: out += path + " str " + hex(v.str) + "\n"; break;
        case V::Arr:
            out += path + " arr " + std::to_string(v.arr.size()) + "\n";
            for (size_t i = 0; i < v.arr.size(); ++i) flat_pico(v.arr[i], path + "[" + std::to_string(i) + "]", out);
            break;
        case V::Obj:
            out += path + " obj " + std::to_string(v.obj.size()) + "\n";
            for (const auto& [k, m] : v.obj) flat_pico(m, path + "." + hex(k), out);
            break;
    }
}

int emit(const std::string& s, char* out, size_t cap) {
    if (s.size() + 1 > cap) return -1;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return 1;
}

}  // namespace

extern "C" {

/* 1: parsed, canonical tree in out; 0: jsoncpp refuses the text; -1: out too small. */
int jc_ref_flatten(const char* text, size_t n, char* out, size_t cap) {
    Json::Value root;
    try {
        std::istringstream in(std::string(text, n));
        in >> root;
    } catch (const std::exception&) {
        return 0;
    }
    std::string s;
    flat_ref(root, "", s);
    return emit(s, out, cap);
}

/* The same with the library's reader. */
int jc_pico_flatten(const char* text, size_t n, char* out, size_t cap) {
    pico_json::Value root;
    const std::string t(text, n);
    try {
        root = pico_json::Reader(t, "").parse();
    } catch (const pico_json::Error&) {
        return 0;
    }
    std::string s;
    flat_pico(root, "", s);
    return emit(s, out, cap);
}

}  // extern "C"
