// TEST INFRASTRUCTURE -- builds the reference's own INI library (external/include/mini/ini.h,
// mINI 0.9.7, header-only) into oracle/_ref/libini_ref.so so the tests can check the settings
// writer and reader against the parser and generator the reference saves its settings with
// (source/main.cpp:4238-4248, 4300-4389). The header is read from /root/reference/external/include
// at build time; nothing of it is copied into this repository.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "mini/ini.h"

namespace {
bool write_file(const char* path, const char* text) {
    std::ofstream f(path, std::ios::binary);
    f << text;
    return (bool)f;
}
std::string read_file(const char* path) {
    std::ifstream f(path, std::ios::binary);
    return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
int copy_out(const std::string& s, char* out, size_t cap) {
    if (s.size() + 1 > cap) return -1;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return (int)s.size();
}
}  // namespace

// Parse `text` with mINI's reader and print it again with generate() (no pretty print), as the
// reference writes its settings file; `tmp_in` / `tmp_out` are scratch file paths.
extern "C" int ref_ini_regenerate(const char* text, const char* tmp_in, const char* tmp_out, char* out, size_t cap) {
    if (!write_file(tmp_in, text)) return -2;
    mINI::INIFile in(tmp_in);
    mINI::INIStructure ini;
    if (!in.read(ini)) return -3;
    mINI::INIFile outf(tmp_out);
    if (!outf.generate(ini)) return -4;
    return copy_out(read_file(tmp_out), out, cap);
}

// The value mINI's reader gives for [section] key ("" when absent), as ini.get(s).get(k).
extern "C" int ref_ini_get(const char* text, const char* tmp_in, const char* section, const char* key, char* out,
                           size_t cap) {
    if (!write_file(tmp_in, text)) return -2;
    mINI::INIFile in(tmp_in);
    mINI::INIStructure ini;
    if (!in.read(ini)) return -3;
    return copy_out(ini.get(section).get(key), out, cap);
}
