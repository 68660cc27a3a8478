// Host driver of the grammar identity for tests/test_identity.py:
//   identity_driver sha            SHA-256 of standard input
//   identity_driver grammar FILE   content hash of a grammar document
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <string>

#include "identity.hpp"

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "sha") {
    const std::string in((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
    std::cout << gagp_amd::sha256_hex(in) << "\n";
    return 0;
  }
  if (mode == "grammar" && argc == 3) {
    std::ifstream f(argv[2]);
    std::stringstream text;
    text << f.rdbuf();
    std::cout << gagp_amd::grammar_content_hash(gagp_amd::json::parse(text.str())) << "\n";
    return 0;
  }
  std::cerr << "usage: identity_driver sha | grammar FILE\n";
  return 2;
}
