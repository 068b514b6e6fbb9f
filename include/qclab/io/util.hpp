// String helpers of the QASM reader (reference include/qclab/io/util.hpp,
// src/io/util.cpp).
#pragma once

#include <string>
#include <vector>

namespace qclab::io {

template <typename T> T read_value(const std::string& str);
template <> inline int read_value(const std::string& str) { return std::stoi(str); }
template <> inline float read_value(const std::string& str) { return std::stof(str); }
template <> inline double read_value(const std::string& str) { return std::stod(str); }
template <> inline std::string read_value(const std::string& str) { return str; }

/// Returns the part of `str` before the first character of `delimiter` and
/// removes it and that character from `str`.  No delimiter: returns the whole
/// string and leaves `str` empty.  Delimiter first: returns "" and drops it.
inline std::string left_of(std::string& str, const std::string delimiter) {
  const auto pos = str.find_first_of(delimiter);
  if (pos == std::string::npos) {
    std::string all;
    all.swap(str);
    return all;
  }
  std::string left = str.substr(0, pos);
  str.erase(0, pos + 1);
  return left;
}

/// Splits `str` at `delimiter` (consumes it); empty pieces are dropped.
template <typename T>
std::vector<T> split(std::string& str, const std::string delimiter) {
  std::vector<T> vec;
  while (!str.empty()) {
    const std::string piece = left_of(str, delimiter);
    if (!piece.empty()) vec.push_back(read_value<T>(piece));
  }
  return vec;
}

}  // namespace qclab::io
