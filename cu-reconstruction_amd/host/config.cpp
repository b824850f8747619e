// config.cpp -- the flat YAML parameter file of the reference's drivers
// (modules/src/config.cpp: an OpenCV FileStorage in YAML mode, read with
// get<T>(key, default)).  Only the subset those files use is parsed: an
// optional "%YAML:1.0" line, "key: value" lines, '#' comments and quoted or
// bare scalars.
#include "recon/tsdf_surface.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace cuphoto {

namespace {

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// A '#' starts a comment unless it is inside quotes.
std::string strip_comment(const std::string& s) {
    char quote = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (quote) {
            if (c == quote) quote = 0;
        } else if (c == '"' || c == '\'') {
            quote = c;
        } else if (c == '#') {
            return s.substr(0, i);
        }
    }
    return s;
}

}  // namespace

bool Config::parse(const std::string& text) {
    values.clear();
    std::istringstream in(text);
    std::string line;
    int n = 0;
    while (std::getline(in, line)) {
        ++n;
        const std::string t = trim(strip_comment(line));
        if (t.empty() || t[0] == '%' || t == "---") continue;
        const size_t colon = t.find(':');
        if (colon == std::string::npos || colon == 0) {
            std::fprintf(stderr, "[cuphoto] parameter file line %d: expected 'key: value'\n", n);
            return false;
        }
        std::string key = trim(t.substr(0, colon));
        std::string val = trim(t.substr(colon + 1));
        if (val.size() >= 2 && (val.front() == '"' || val.front() == '\'') && val.back() == val.front())
            val = val.substr(1, val.size() - 2);
        values[key] = val;
    }
    return true;
}

bool Config::set_parameter_file(const std::string& filename) {
    std::ifstream f(filename);
    if (!f.is_open()) {
        std::fprintf(stderr, "[cuphoto] parameter file %s does not exist.\n", filename.c_str());
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

template <>
double Config::get<double>(const std::string& key, double def) const {
    const auto it = values.find(key);
    if (it == values.end()) return def;
    char* end = nullptr;
    const double v = std::strtod(it->second.c_str(), &end);
    return end == it->second.c_str() ? def : v;
}

template <>
float Config::get<float>(const std::string& key, float def) const {
    return (float)get<double>(key, (double)def);
}

template <>
int Config::get<int>(const std::string& key, int def) const {
    return (int)get<double>(key, (double)def);
}

template <>
std::string Config::get<std::string>(const std::string& key, std::string def) const {
    const auto it = values.find(key);
    return it == values.end() ? def : it->second;
}

}  // namespace cuphoto
