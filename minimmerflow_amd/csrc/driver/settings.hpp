// settings.hpp -- reader for the reference's settings.xml (main.cpp:73-74).
//
// A small element tree: tags with text or child elements, attributes ignored,
// comments and the <?xml ...?> header skipped.  Lookup by dotted path, as the
// reference's config::root["problem"]["domain"].get<double>("length").
#pragma once

#include <cctype>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mmfdrv {

struct Node {
  std::string name;
  std::string text;
  std::vector<std::unique_ptr<Node>> children;

  const Node *child(const std::string &n) const {
    for (const auto &c : children)
      if (c->name == n) return c.get();
    return nullptr;
  }
  // dotted path lookup, e.g. "problem.domain.length"
  const Node *find(const std::string &path) const {
    const Node *cur = this;
    std::stringstream ss(path);
    std::string part;
    while (cur && std::getline(ss, part, '.')) cur = cur->child(part);
    return cur;
  }
  bool has(const std::string &path) const { return find(path) != nullptr; }
  std::string str(const std::string &path) const {
    const Node *n = find(path);
    if (!n) throw std::runtime_error("settings: missing key " + path);
    return n->text;
  }
  std::string str(const std::string &path, const std::string &dflt) const {
    const Node *n = find(path);
    return n ? n->text : dflt;
  }
  double num(const std::string &path) const { return std::stod(str(path)); }
  double num(const std::string &path, double dflt) const {
    const Node *n = find(path);
    return n ? std::stod(n->text) : dflt;
  }
};

class Parser {
 public:
  explicit Parser(const std::string &s) : s_(s) {}

  std::unique_ptr<Node> parse() {
    skip_misc();
    auto root = element();
    return root;
  }

 private:
  const std::string &s_;
  size_t i_ = 0;

  void ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  bool starts(const char *p) const { return s_.compare(i_, std::char_traits<char>::length(p), p) == 0; }
  void skip_misc() {
    for (;;) {
      ws();
      if (starts("<?")) {
        size_t e = s_.find("?>", i_);
        if (e == std::string::npos) throw std::runtime_error("settings: unterminated header");
        i_ = e + 2;
      } else if (starts("<!--")) {
        size_t e = s_.find("-->", i_);
        if (e == std::string::npos) throw std::runtime_error("settings: unterminated comment");
        i_ = e + 3;
      } else {
        return;
      }
    }
  }
  static std::string trim(const std::string &t) {
    size_t b = 0, e = t.size();
    while (b < e && std::isspace((unsigned char)t[b])) ++b;
    while (e > b && std::isspace((unsigned char)t[e - 1])) --e;
    return t.substr(b, e - b);
  }
  std::unique_ptr<Node> element() {
    if (i_ >= s_.size() || s_[i_] != '<') throw std::runtime_error("settings: expected an element");
    ++i_;
    size_t b = i_;
    while (i_ < s_.size() && !std::isspace((unsigned char)s_[i_]) && s_[i_] != '>' && s_[i_] != '/') ++i_;
    auto n = std::make_unique<Node>();
    n->name = s_.substr(b, i_ - b);
    size_t gt = s_.find('>', i_);
    if (gt == std::string::npos) throw std::runtime_error("settings: unterminated tag " + n->name);
    bool selfclose = gt > 0 && s_[gt - 1] == '/';
    i_ = gt + 1;
    if (selfclose) return n;
    std::string text;
    for (;;) {
      if (i_ >= s_.size()) throw std::runtime_error("settings: unterminated element " + n->name);
      if (starts("<!--")) {
        skip_misc();
        continue;
      }
      if (starts("</")) {
        size_t e = s_.find('>', i_);
        if (e == std::string::npos) throw std::runtime_error("settings: bad closing tag");
        std::string close = trim(s_.substr(i_ + 2, e - i_ - 2));
        if (close != n->name) throw std::runtime_error("settings: <" + n->name + "> closed by </" + close + ">");
        i_ = e + 1;
        break;
      }
      if (s_[i_] == '<') {
        n->children.push_back(element());
      } else {
        text.push_back(s_[i_++]);
      }
    }
    n->text = trim(text);
    return n;
  }
};

inline std::unique_ptr<Node> read_settings(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("settings: cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  std::string content = ss.str();
  Parser p(content);
  return p.parse();
}

}  // namespace mmfdrv
