// cmdopts.hpp -- GOWalker's command-line option semantics (api/cmdopts.hpp:108-246,
// api/configfile.hpp, api/filename.hpp), kept so that the same command lines work:
//   Node2vec <graph.txt> length 40 walkpersource 1 memory 16 blocksize 256 p 0.5 q 2
//            gpu cpu zero skip GOWalker.
// * `name value` pairs are looked up by scanning argv backwards from the second-to-last
//   token (get_option_int/long/float/string, cmdopts.hpp:191-235);
// * a bare flag is true if any token equals its name (get_option_bool, :237-246);
// * `--key=value` tokens and the config file conf/randgraph.cnf (directory from
//   $RANDGRAPH_ROOT, filename.hpp:18-29) fill a key/value map consulted when argv has no
//   pair.  The reference opens but never parses the config file (configfile.hpp:97-117);
//   here `key = value` lines are parsed, and the command line always wins.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>

namespace gwh {

class Options {
public:
    void set_argc(int argc, const char** argv)
    {
        argc_ = argc;
        argv_ = argv;
        load_config_file();
        for (int i = 1; i < argc; ++i) {
            std::string a(argv[i]);
            if (a.compare(0, 2, "--") != 0) continue;
            size_t eq = a.find('=');
            if (eq == std::string::npos) continue;
            conf_[a.substr(2, eq - 2)] = a.substr(eq + 1);
        }
    }

    const char* find_pair(const char* name) const
    {
        for (int i = argc_ - 2; i >= 0; --i)
            if (std::strcmp(argv_[i], name) == 0) return argv_[i + 1];
        return nullptr;
    }

    int get_int(const char* name, int def) const
    {
        if (const char* v = find_pair(name)) return std::atoi(v);
        auto it = conf_.find(name);
        return it == conf_.end() ? def : std::atoi(it->second.c_str());
    }

    unsigned long long get_long(const char* name, unsigned long long def) const
    {
        if (const char* v = find_pair(name)) return std::strtoull(v, nullptr, 10);
        auto it = conf_.find(name);
        return it == conf_.end() ? def : std::strtoull(it->second.c_str(), nullptr, 10);
    }

    float get_float(const char* name, float def) const
    {
        if (const char* v = find_pair(name)) return (float)std::atof(v);
        auto it = conf_.find(name);
        return it == conf_.end() ? def : (float)std::atof(it->second.c_str());
    }

    std::string get_string(const char* name, const std::string& def) const
    {
        if (const char* v = find_pair(name)) return v;
        auto it = conf_.find(name);
        return it == conf_.end() ? def : it->second;
    }

    bool get_bool(const char* name) const
    {
        for (int i = argc_ - 1; i >= 0; --i)
            if (std::strcmp(argv_[i], name) == 0) return true;
        return false;
    }

    static std::string config_path()
    {
        const char* root = std::getenv("RANDGRAPH_ROOT");
        std::string dir = root ? root : ".";
        return dir + "/conf/randgraph.cnf";
    }

private:
    void load_config_file()
    {
        std::ifstream f(config_path());
        std::string line;
        while (std::getline(f, line)) {
            size_t h = line.find('#');
            if (h != std::string::npos) line.resize(h);
            size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            auto trim = [](std::string s) {
                size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
                return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
            };
            std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
            if (!k.empty()) conf_[k] = v;
        }
    }

    int argc_ = 0;
    const char** argv_ = nullptr;
    std::map<std::string, std::string> conf_;
};

}  // namespace gwh
