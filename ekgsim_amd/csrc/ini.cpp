// ini.cpp -- see ini.hpp for the accepted syntax.
#include "ini.hpp"

#include <fstream>

namespace ekgapp {

namespace {

bool blank(char c) { return c == ' ' || c == '\t'; }

// Split one line into (name, value): both trimmed of blanks; the value is
// everything after the first '='.  Mirrors the character scan of
// Ini.cpp:357-381 (a blank only advances the start while nothing else has
// been seen; the end is the last non-blank character).
void split_line(const std::string& in, std::string& name, std::string& val, bool& has_eq)
{
    size_t first = 0, last = 0;
    bool eq = false;
    bool any = false;
    for (size_t i = 0; i < in.size(); ++i) {
        if (in[i] == '=' && !eq) {
            name = any ? in.substr(first, last - first + 1) : std::string();
            eq = true;
            first = i + 1;
            any = false;
            continue;
        }
        if (blank(in[i])) {
            if (i == first) ++first;
        } else {
            last = i;
            any = true;
        }
    }
    if (eq) {
        val = (any && last >= first) ? in.substr(first, last - first + 1) : std::string();
    } else {
        name = any ? in.substr(first, last - first + 1) : std::string();
        val.clear();
    }
    has_eq = eq;
}

// "[ name ]" -> "name" (Ini.cpp, Section::normalizeName).
std::string normalize_section(const std::string& n)
{
    size_t a = 0;
    bool br = false;
    while (a < n.size()) {
        if (blank(n[a])) {
            ++a;
        } else if (!br && n[a] == '[') {
            br = true;
            ++a;
        } else {
            break;
        }
    }
    size_t b = n.size();
    br = false;
    while (b > a) {
        if (blank(n[b - 1])) {
            --b;
        } else if (!br && n[b - 1] == ']') {
            br = true;
            --b;
        } else {
            break;
        }
    }
    return n.substr(a, b - a);
}

}  // namespace

IniFile::IniFile(const std::string& path)
{
    secs_.push_back(Sec{"default", {}});
    std::ifstream f(path.c_str());
    if (!f.is_open()) return;
    found_ = true;
    std::string line, name, val;
    size_t active = 0;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        bool eq = false;
        split_line(line, name, val, eq);
        if (!name.empty() && name.front() == '[' && name.back() == ']') {
            const std::string sn = normalize_section(name);
            active = secs_.size();
            for (size_t s = 0; s < secs_.size(); ++s)
                if (secs_[s].name == sn) active = s;
            if (active == secs_.size()) secs_.push_back(Sec{sn, {}});
        } else {
            secs_[active].vars.push_back(Var{name, val});
            empty_ = false;
        }
    }
}

int IniFile::section(const std::string& name) const
{
    for (size_t s = 1; s < secs_.size(); ++s)
        if (secs_[s].name == name) return (int)s;
    return -1;
}

bool IniFile::get(std::string& out, const std::string& key, int sec) const
{
    if (sec < 0 || sec >= (int)secs_.size()) return false;
    const auto& vars = secs_[sec].vars;
    for (size_t i = vars.size(); i > 0; --i)
        if (vars[i - 1].name == key) {
            out = vars[i - 1].value;
            return true;
        }
    return false;
}

}  // namespace ekgapp
