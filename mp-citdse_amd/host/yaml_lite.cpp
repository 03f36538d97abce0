#include "yaml_lite.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace mpci {

namespace {

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

// Drop a trailing comment that is not inside quotes.
std::string strip_comment(const std::string &s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (q) {
      if (c == q) q = 0;
    } else if (c == '"' || c == '\'') {
      q = c;
    } else if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) {
      return s.substr(0, i);
    }
  }
  return s;
}

std::string unquote(const std::string &s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') ||
                        (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

}  // namespace

YamlDoc YamlDoc::load_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw YamlError("cannot open YAML file " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

YamlDoc YamlDoc::parse(const std::string &text) {
  YamlDoc d;
  std::istringstream in(text);
  std::string line, section;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    std::string body = strip_comment(line);
    if (trim(body).empty()) continue;
    const bool indented = std::isspace((unsigned char)body[0]);
    std::string t = trim(body);
    size_t colon = t.find(':');
    if (colon == std::string::npos)
      throw YamlError("line " + std::to_string(lineno) + ": expected key:");
    std::string key = trim(t.substr(0, colon));
    std::string val = trim(t.substr(colon + 1));
    if (!indented) {
      section = key;
      d.v_[section];
      if (!val.empty()) d.v_[""][key] = unquote(val);
      continue;
    }
    if (section.empty())
      throw YamlError("line " + std::to_string(lineno) +
                      ": indented key outside a section");
    d.v_[section][key] = (val.size() && val[0] == '[') ? val : unquote(val);
  }
  return d;
}

bool YamlDoc::has(const std::string &section, const std::string &key) const {
  auto it = v_.find(section);
  return it != v_.end() && it->second.count(key);
}

const std::string &YamlDoc::raw(const std::string &section,
                                const std::string &key) const {
  auto it = v_.find(section);
  if (it == v_.end() || !it->second.count(key))
    throw YamlError("missing setting " + section + "." + key);
  return it->second.at(key);
}

std::string YamlDoc::str(const std::string &section,
                         const std::string &key) const {
  return raw(section, key);
}

int YamlDoc::as_int(const std::string &section, const std::string &key) const {
  const std::string &s = raw(section, key);
  char *end = nullptr;
  errno = 0;
  long v = std::strtol(s.c_str(), &end, 10);
  if (s.empty() || *end != '\0' || errno)
    throw YamlError("bad conversion of " + section + "." + key + " to int");
  return (int)v;
}

double YamlDoc::as_double(const std::string &section,
                          const std::string &key) const {
  const std::string &s = raw(section, key);
  char *end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (s.empty() || *end != '\0')
    throw YamlError("bad conversion of " + section + "." + key + " to double");
  return v;
}

char YamlDoc::as_char(const std::string &section,
                      const std::string &key) const {
  const std::string &s = raw(section, key);
  if (s.size() != 1)
    throw YamlError("bad conversion of " + section + "." + key + " to char");
  return s[0];
}

bool YamlDoc::as_bool(const std::string &section,
                      const std::string &key) const {
  const std::string &s = raw(section, key);
  if (s == "true" || s == "True" || s == "yes" || s == "1") return true;
  if (s == "false" || s == "False" || s == "no" || s == "0") return false;
  throw YamlError("bad conversion of " + section + "." + key + " to bool");
}

std::vector<int> YamlDoc::as_int_list(const std::string &section,
                                      const std::string &key) const {
  std::string s = trim(raw(section, key));
  if (s.size() < 2 || s.front() != '[' || s.back() != ']')
    throw YamlError(section + "." + key + " is not a [..] list");
  std::vector<int> out;
  std::string inner = s.substr(1, s.size() - 2);
  std::stringstream ss(inner);
  std::string item;
  while (std::getline(ss, item, ',')) {
    item = trim(item);
    if (item.empty()) continue;
    char *end = nullptr;
    long v = std::strtol(item.c_str(), &end, 10);
    if (*end != '\0')
      throw YamlError("bad integer '" + item + "' in " + section + "." + key);
    out.push_back((int)v);
  }
  return out;
}

}  // namespace mpci
