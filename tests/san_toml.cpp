// Host-side check of the TOML reader under AddressSanitizer / UndefinedBehaviorSanitizer:
// every job file passed on the command line parses, and malformed texts (unterminated
// strings, arrays and headers, missing '=', deep nesting, out-of-range integers) are
// refused with an exception instead of reading out of bounds or looping.
#include <iostream>
#include <string>

#include "../icarat-contact_amd/host/toml_lite.hpp"

using icarat::toml_lite::parse;
using icarat::toml_lite::parse_string;

int main(int argc, char **argv) {
  int bad = 0;
  for (int i = 1; i < argc; ++i) {
    const auto t = parse(argv[i]);
    const auto &div = t.at("mesh", "division").as_array();
    if (div.size() != 3) { std::cerr << argv[i] << ": division\n"; ++bad; }
    (void)div[0].as_int();
  }
  const char *broken[] = {
      "[mesh\ndimension = 2\n",
      "[mesh]\ntype = \"quad4\n",
      "[mesh]\ndivision = [1, 2, 3\n",
      "[mesh]\ndivision = [1, 2,\n",
      "[mesh]\ndimension\n",
      "[mesh]\nx = [[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[[1]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]]\n",
      "[mesh]\ndivision = [1e300, 1, 1]\n",
      "[mesh]\nx = \n",
      "[mesh]\nx = [1 2]\n",
  };
  for (const char *s : broken) {
    try {
      const auto t = parse_string(s);
      if (t.has("mesh", "division")) (void)t.at("mesh", "division").as_array().at(0).as_int();
      std::cerr << "accepted a malformed text: " << s << "\n";
      ++bad;
    } catch (const std::exception &) {
    }
  }
  if (bad) return 1;
  std::cout << "toml reader: ok\n";
  return 0;
}
