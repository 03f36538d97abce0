// Reader for the YAML subset of MP-CITDSE's input files (inp/*.yaml):
// top-level sections holding `key: value` pairs, scalars (numbers, bare or
// quoted strings), flow lists `[a, b, c]` and `#` comments.  It stands in for
// the yaml-cpp calls of tdrd::readConfig (src/tdse/td_read.cpp:9-43), with
// the same conversions: as<int>, as<double>, as<char> (first character of
// the string), as<std::string>, as<std::vector<int>>.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpci {

class YamlError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class YamlDoc {
 public:
  static YamlDoc load_file(const std::string &path);
  static YamlDoc parse(const std::string &text);

  bool has(const std::string &section, const std::string &key) const;
  std::string str(const std::string &section, const std::string &key) const;
  int as_int(const std::string &section, const std::string &key) const;
  double as_double(const std::string &section, const std::string &key) const;
  char as_char(const std::string &section, const std::string &key) const;
  bool as_bool(const std::string &section, const std::string &key) const;
  std::vector<int> as_int_list(const std::string &section,
                               const std::string &key) const;

 private:
  // raw scalar text (quotes removed) or "[...]" for flow lists
  std::map<std::string, std::map<std::string, std::string>> v_;
  const std::string &raw(const std::string &section,
                         const std::string &key) const;
};

}  // namespace mpci
