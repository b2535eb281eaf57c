// registry.hpp -- reader for MARQOV's .ini configuration files.
//
// Same file format and lookup rules as the reference's RegistryDB
// (src/libmarqov/util/registry.h:254-310, registry.cpp:27-251), so the
// reference's config/*.ini files load unchanged:
//   * every file of the configuration directory whose name ends in the suffix
//     is read; lines are trimmed of blanks and tabs, empty lines and lines
//     starting with '#' are dropped;
//   * a file must contain an "[END]" line, else it is not a config file;
//   * a line containing '[' opens a block named by the text between the first
//     '[' and the last ']'; the block runs to the next such line;
//   * "key = value" lines split at the first '=', both sides trimmed;
//   * scalars convert with stream extraction; vectors split at any of ';', ' ',
//     ',' (registry.h:399-423); booleans are true iff the text is TRUE in any case.
// Missing files, blocks and keys throw RegistryError naming what is missing
// (the reference throws its Registry_*_not_found exceptions, registry.h:45-112).
#pragma once

#include <dirent.h>

#include <algorithm>
#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace marqov_amd {

class RegistryError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

namespace detail {

inline std::string trim(const std::string& s)
{
    const auto b = s.find_first_not_of(" \t\r");
    if (b == std::string::npos) return "";
    const auto e = s.find_last_not_of(" \t\r");
    return s.substr(b, e - b + 1);
}

template <class T>
struct Convert {
    static T from(const std::string& text, const std::string& where)
    {
        std::istringstream in(text);
        T v{};
        in >> v;
        if (in.fail()) throw RegistryError("cannot convert '" + text + "' for " + where);
        return v;
    }
};

template <>
struct Convert<std::string> {
    static std::string from(const std::string& text, const std::string&) { return text; }
};

template <>
struct Convert<bool> {
    static bool from(const std::string& text, const std::string&)
    {
        std::string u(text);
        std::transform(u.begin(), u.end(), u.begin(), [](unsigned char c) { return (char)std::toupper(c); });
        return u == "TRUE";
    }
};

template <class T>
struct Convert<std::vector<T>> {
    static std::vector<T> from(const std::string& text, const std::string& where)
    {
        static const char* delim = "; ,";
        std::vector<T> out;
        std::size_t pos = text.find_first_not_of(delim);
        while (pos != std::string::npos) {
            const std::size_t end = text.find_first_of(delim, pos);
            out.push_back(Convert<T>::from(text.substr(pos, end == std::string::npos ? std::string::npos : end - pos),
                                           where));
            pos = end == std::string::npos ? end : text.find_first_not_of(delim, end);
        }
        if (out.empty()) throw RegistryError("empty list for " + where);
        return out;
    }
};

}  // namespace detail

class Registry {
public:
    using Block = std::map<std::string, std::string>;
    using File = std::map<std::string, Block>;

    Registry() = default;
    // Reads every file in `dir` whose name ends in `suffix` (RegistryDB::init).
    explicit Registry(const std::string& dir, const std::string& suffix = ".ini") { init(dir, suffix); }

    void init(const std::string& dir, const std::string& suffix = ".ini")
    {
        DIR* d = opendir(dir.c_str());
        if (!d) throw RegistryError("Configuration Directory not found or empty: " + dir);
        std::vector<std::string> names;
        while (dirent* e = readdir(d)) {
            const std::string n(e->d_name);
            if (n == "." || n == "..") continue;
            if (n.size() > suffix.size() && n.compare(n.size() - suffix.size(), suffix.size(), suffix) == 0)
                names.push_back(n);
        }
        closedir(d);
        std::sort(names.begin(), names.end());
        for (const auto& n : names) files_[n] = parse_file(dir + "/" + n);
    }

    // Parses one file's text (the body of cfile::cfile, registry.cpp:89-123).
    static File parse_text(const std::string& text, const std::string& name)
    {
        std::vector<std::string> lines;
        std::istringstream in(text);
        std::string line;
        while (std::getline(in, line)) {
            line = detail::trim(line);
            if (!line.empty() && line[0] != '#') lines.push_back(line);
        }
        if (std::find(lines.begin(), lines.end(), std::string("[END]")) == lines.end())
            throw RegistryError("File not recognized as Config File! " + name);
        File f;
        std::size_t k = 0;
        auto next_bracket = [&](std::size_t from) {
            for (std::size_t i = from; i < lines.size(); ++i)
                if (lines[i].find('[') != std::string::npos) return i;
            return lines.size();
        };
        k = next_bracket(0);
        while (k < lines.size()) {
            const std::string& hdr = lines[k];
            const auto first = hdr.find('[');
            const auto last = hdr.rfind(']');
            if (last == std::string::npos || last < first)
                throw RegistryError("Error in Parsing ConfigFile!! Closing Bracket not found in " + name);
            const std::string bname = detail::trim(hdr.substr(first + 1, last - first - 1));
            if (bname == "END") break;
            const std::size_t nb = next_bracket(k + 1);
            if (f.count(bname)) {  // a repeated block name keeps its first definition (map insert)
                k = nb;
                continue;
            }
            Block& b = f[bname];
            for (std::size_t i = k + 1; i < nb; ++i) {
                const auto eq = lines[i].find('=');
                const std::string key = detail::trim(lines[i].substr(0, eq));
                const std::string val = eq == std::string::npos ? std::string() : detail::trim(lines[i].substr(eq + 1));
                b[key] = val;
            }
            k = nb;
        }
        return f;
    }

    static File parse_file(const std::string& path)
    {
        std::ifstream in(path);
        if (!in) throw RegistryError("Couldn't open Config File : " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        return parse_text(ss.str(), path);
    }

    void add(const std::string& file, const File& content) { files_[file] = content; }
    bool has_file(const std::string& file) const { return files_.count(file) != 0; }
    bool has(const std::string& file, const std::string& block, const std::string& key) const
    {
        auto f = files_.find(file);
        if (f == files_.end()) return false;
        auto b = f->second.find(block);
        return b != f->second.end() && b->second.count(key) != 0;
    }

    template <class T>
    T get(const std::string& file, const std::string& block, const std::string& key) const
    {
        auto f = files_.find(file);
        if (f == files_.end()) throw RegistryError("File Key not found :" + file);
        auto b = f->second.find(block);
        if (b == f->second.end()) throw RegistryError("Block not found: " + block + " (in " + file + ")");
        auto v = b->second.find(key);
        if (v == b->second.end()) throw RegistryError("Block Key not found: " + key + " (in [" + block + "] of " + file + ")");
        return detail::Convert<T>::from(v->second, file + ":[" + block + "]:" + key);
    }

    template <class T>
    T get_or(const std::string& file, const std::string& block, const std::string& key, const T& dflt) const
    {
        return has(file, block, key) ? get<T>(file, block, key) : dflt;
    }

private:
    std::map<std::string, File> files_;
};

}  // namespace marqov_amd
