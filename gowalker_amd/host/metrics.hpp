// metrics.hpp -- run metrics in the format of GOWalker's console and file reporters
// (metrics/reps/file_reporter.hpp:56-90, appended to randgraph_metrics.txt): a "[name]"
// header, then "name.key=value" lines, with .count/.min/.max/.avg for numbers, so that
// scripts grepping keys such as "0_Total_time" keep working.
#pragma once
#include <cstdio>
#include <map>
#include <string>

namespace gwh {

class Metrics {
public:
    explicit Metrics(std::string name) : name_(std::move(name)) {}

    void set(const std::string& key, double v, bool integer = false)
    {
        Entry& e = m_[key];
        e.integer = integer;
        e.value = v;
        e.sum += v;
        if (!e.count || v < e.min) e.min = v;
        if (!e.count || v > e.max) e.max = v;
        e.count++;
    }
    void set_string(const std::string& key, const std::string& v) { s_[key] = v; }

    void write(FILE* f) const
    {
        std::fprintf(f, "[%s]\n", name_.c_str());
        for (auto& kv : m_) {
            const Entry& e = kv.second;
            const char* k = kv.first.c_str();
            if (e.integer) {
                std::fprintf(f, "%s.%s=%lld\n", name_.c_str(), k, (long long)e.value);
                std::fprintf(f, "%s.%s.count=%llu\n", name_.c_str(), k, (unsigned long long)e.count);
                std::fprintf(f, "%s.%s.min=%lld\n", name_.c_str(), k, (long long)e.min);
                std::fprintf(f, "%s.%s.max=%lld\n", name_.c_str(), k, (long long)e.max);
            } else {
                std::fprintf(f, "%s.%s=%lf\n", name_.c_str(), k, e.value);
                std::fprintf(f, "%s.%s.count=%llu\n", name_.c_str(), k, (unsigned long long)e.count);
                std::fprintf(f, "%s.%s.min=%lf\n", name_.c_str(), k, e.min);
                std::fprintf(f, "%s.%s.max=%lf\n", name_.c_str(), k, e.max);
            }
            std::fprintf(f, "%s.%s.avg=%lf\n", name_.c_str(), k, e.sum / (double)e.count);
        }
        for (auto& kv : s_) std::fprintf(f, "%s.%s=%s\n", name_.c_str(), kv.first.c_str(), kv.second.c_str());
    }

    void report(const std::string& file) const
    {
        write(stdout);
        if (FILE* f = std::fopen(file.c_str(), "a")) {
            write(f);
            std::fclose(f);
        } else {
            std::fprintf(stderr, "cannot append metrics to %s\n", file.c_str());
        }
    }

private:
    struct Entry {
        bool integer = false;
        double value = 0, sum = 0, min = 0, max = 0;
        unsigned long long count = 0;
    };
    std::string name_;
    std::map<std::string, Entry> m_;
    std::map<std::string, std::string> s_;
};

}  // namespace gwh
