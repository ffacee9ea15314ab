// Command-line parsing with the reference's rules (lib/tools/arg_parser.h:21-61):
// "-key value" sets key to value when the next argument does not start with
// '-', otherwise "-key" is a switch with an empty value; other arguments are
// positional.  Values are converted with stream extraction.
#pragma once

#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace kc {

class ArgParser {
 public:
  ArgParser(int argc, char **argv) {
    int pos = 1;
    while (pos < argc) {
      if (argv[pos][0] == '-') {
        std::string key(argv[pos] + 1);
        ++pos;
        if (pos < argc && argv[pos][0] != '-')
          named_[key] = argv[pos++];
        else
          named_[key] = "";
      } else {
        positional_.emplace_back(argv[pos++]);
      }
    }
  }

  template <class T>
  T Get(const std::string &key, const T &fallback) const {
    auto it = named_.find(key);
    if (it == named_.end()) return fallback;
    T v{};
    std::istringstream s(it->second);
    s >> v;
    return v;
  }

  bool IsSet(const std::string &key) const { return named_.count(key) != 0; }
  const std::vector<std::string> &Positional() const { return positional_; }

 private:
  std::unordered_map<std::string, std::string> named_;
  std::vector<std::string> positional_;
};

template <>
inline std::string ArgParser::Get<std::string>(const std::string &key, const std::string &fallback) const {
  auto it = named_.find(key);
  if (it == named_.end()) return fallback;
  std::istringstream s(it->second);
  std::string v;
  s >> v;
  return v;
}

}  // namespace kc
