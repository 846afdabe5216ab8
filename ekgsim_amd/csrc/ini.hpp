// ini.hpp -- reader for simulator.ini with the reference's parsing rules
// (copyOfLibs/Ini.cpp:357-381, Ini.h:142-299):
//   - a line is "name = value" or "name"; spaces and tabs around both are
//     trimmed; everything after the first '=' is the value (no inline
//     comments: "key = 3 ; note" has the value "3 ; note");
//   - "[section]" lines open a section (name trimmed of blanks and one pair
//     of brackets); repeated sections merge;
//   - within a section the last occurrence of a key wins;
//   - typed reads use stream extraction (">>") on the value; array reads
//     split on any of " ;,\t".
#pragma once

#include <sstream>
#include <string>
#include <vector>

namespace ekgapp {

class IniFile {
public:
    explicit IniFile(const std::string& path);

    bool found() const { return found_; }
    bool empty() const { return empty_; }
    // section index, -1 when absent (the unnamed leading section is 0)
    int section(const std::string& name) const;

    bool get(std::string& out, const std::string& key, int sec) const;
    template <class T>
    bool get(T& out, const std::string& key, int sec) const
    {
        std::string v;
        if (!get(v, key, sec)) return false;
        std::istringstream s(v);
        s >> out;
        return true;
    }
    // Array of up to `limit` elements (0 = no limit), separated by any of
    // " ;,\t" (Ini.h:267-299).  Returns false when the key is absent.
    template <class T>
    bool get_array(std::vector<T>& out, const std::string& key, int sec, unsigned limit = 0) const
    {
        std::string v;
        if (!get(v, key, sec)) return false;
        std::istringstream s(v);
        unsigned left = limit;
        while (s) {
            T x{};
            s >> x;
            out.push_back(x);
            bool removed = false;
            while (s.good()) {
                const int ch = s.peek();
                if (ch == ' ' || ch == ';' || ch == ',' || ch == '\t') {
                    s.ignore();
                    removed = true;
                } else {
                    break;
                }
            }
            if (!removed) break;
            if (left != 0 && --left == 0) break;
        }
        return true;
    }

private:
    struct Var {
        std::string name, value;
    };
    struct Sec {
        std::string name;
        std::vector<Var> vars;  // in file order; lookups scan from the back
    };
    std::vector<Sec> secs_;
    bool found_ = false, empty_ = true;
};

}  // namespace ekgapp
