// honei/util/configuration.hh -- key/value run-time configuration
// (util/configuration.cc:95-230): `ell::threads` selects the ELL lane
// layout of newly built matrices, as in the reference.
#pragma once

#include <cstdlib>
#include <map>
#include <mutex>
#include <string>

namespace honei {

class Configuration {
  public:
    static Configuration *instance() {
        static Configuration c;
        return &c;
    }
    int get_value(const std::string &key, int def) {
        std::lock_guard<std::mutex> l(_m);
        auto it = _ints.find(key);
        return it == _ints.end() ? def : it->second;
    }
    unsigned long get_value(const std::string &key, unsigned long def) {
        return (unsigned long)get_value(key, (int)def);
    }
    void set_value(const std::string &key, int value) {
        std::lock_guard<std::mutex> l(_m);
        _ints[key] = value;
    }

  private:
    Configuration() {
        if (const char *t = std::getenv("HONEI_ELL_THREADS")) _ints["ell::threads"] = std::atoi(t);
    }
    std::mutex _m;
    std::map<std::string, int> _ints;
};

}  // namespace honei
