// Exception type of the drop-in interface, with the reference's key/value
// entries (StormError.hpp:13-57) and its `STORMError` alias (:61).
#pragma once

#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace STORM {

class StormError : public std::runtime_error {
 public:
  explicit StormError(const std::string &msg) : std::runtime_error(msg), msg_(msg) {}

  template <typename V>
  void addEntry(const std::string &name, const V &value) {
    std::ostringstream oss;
    oss << value;
    entries_.emplace_back(name, oss.str());
    full_ = msg_;
    for (const auto &e : entries_) full_ += "\n  " + e.first + ": " + e.second;
  }

  const std::string &getErrorMessage() const { return msg_; }
  const std::vector<std::pair<std::string, std::string>> &getEntries() const { return entries_; }
  const char *what() const noexcept override { return full_.empty() ? msg_.c_str() : full_.c_str(); }

 private:
  std::string msg_, full_;
  std::vector<std::pair<std::string, std::string>> entries_;
};

inline void reportError(const StormError &eo, std::ostream &os = std::cout) {
  os << eo.getErrorMessage() << std::endl;
  for (const auto &e : eo.getEntries()) os << "  " << e.first << ": " << e.second << std::endl;
}

}  // namespace STORM

using STORMError = STORM::StormError;
