// multiply: command-line drop-in for the reference's `multiply` tool
// (multiply.cpp:44-243): reads A and B from .npy files, computes
// C = A^T B (A is stored n x m, i.e. transposed) with dense FP64 GEMM or the
// compressed FFT / FWHT product on the GPU, and writes C as .npy.
//
//   multiply <A.npy> <B.npy> <C.npy> [-a gemm|fft|fwht] [-b B] [-d D] [-f] [-v...]
//
// Same arguments, checks, messages and exit codes as the reference: the
// output must not exist unless -f; -b (a positive power of two) and -d
// (positive) are required for fft/fwht; both inputs must be 2-D with 8-byte
// words, B must have as many rows as A; Fortran-order inputs are converted
// to C order; the compressed products use multiply-add-shift hashes with a
// clock-derived seed (multiply.cpp:206-218).  Additions: --seed S fixes the
// hash seed and --hash NAME picks another family.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/compmatmul_amd.h"
#include "npy.hpp"

namespace {

void print_matrix(FILE* f, const double* M, int rows, int cols) {
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols; ++j) std::fprintf(f, "%s%2.0f", j > 0 ? " " : "", M[size_t(i) * cols + j]);
    std::fprintf(f, "\n");
  }
}

// rows x cols stored row-major, printed transposed
void print_matrix_t(FILE* f, const double* M, int rows, int cols) {
  for (int j = 0; j < cols; ++j) {
    for (int i = 0; i < rows; ++i) std::fprintf(f, "%s%2.0f", i > 0 ? " " : "", M[size_t(i) * cols + j]);
    std::fprintf(f, "\n");
  }
}

// a Fortran-order rows x cols array in place to C order
void fortran_to_c(int rows, int cols, double* a) {
  std::vector<double> t(size_t(rows) * cols);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) t[size_t(i) * cols + j] = a[size_t(j) * rows + i];
  std::memcpy(a, t.data(), t.size() * sizeof(double));
}

struct Args {
  std::vector<std::string> files;
  bool force = false;
  int b = -1, d = -1;
  bool b_set = false, d_set = false;
  std::string algo = "gemm";
  int verbose = 0;
  bool seed_set = false;
  uint64_t seed = 0;
  int family = CM_MULTIPLY_ADD_SHIFT;
};

struct ArgError : std::runtime_error {
  std::string id;
  ArgError(const std::string& msg, const std::string& arg) : std::runtime_error(msg), id(arg) {}
};

int parse_int(const std::string& v, const std::string& id) {
  size_t used = 0;
  int x = 0;
  try {
    x = std::stoi(v, &used);
  } catch (const std::exception&) {
    used = 0;
  }
  if (used == 0 || used != v.size()) throw ArgError("Couldn't read argument value from string '" + v + "'", id);
  return x;
}

void usage(const char* prog) {
  std::cout << "USAGE: \n\n   " << prog
            << " [-v] ... [-a <gemm|fft|fwht>] [-d <int>] [-b <int>] [-f] [--seed <uint64>] "
               "[--hash <name>] [--] [--version] [-h] <A.npy> <B.npy> <C.npy>\n\n"
               "A simple matrix multiplication program. Computes C=A^TB: the left-hand operand is "
               "stored transposed (n x m), the right-hand one as n x p.\n";
}

Args parse(int argc, char** argv) {
  Args a;
  bool only_files = false;
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    auto value = [&](const std::string& id) -> std::string {
      if (i + 1 >= argc) throw ArgError("Missing a value for this argument!", id);
      return argv[++i];
    };
    if (only_files || s.empty() || s[0] != '-' || s == "-") {
      a.files.push_back(s);
    } else if (s == "--") {
      only_files = true;
    } else if (s == "-h" || s == "--help") {
      usage(argv[0]);
      std::exit(EXIT_SUCCESS);
    } else if (s == "--version") {
      std::cout << "\n" << argv[0] << "  version: compmatmul_amd\n\n";
      std::exit(EXIT_SUCCESS);
    } else if (s == "-b" || s == "--buckets") {
      a.b = parse_int(value("-b (--buckets)"), "-b (--buckets)");
      a.b_set = true;
    } else if (s == "-d" || s == "--sketches") {
      a.d = parse_int(value("-d (--sketches)"), "-d (--sketches)");
      a.d_set = true;
    } else if (s == "-a" || s == "--algorithm") {
      a.algo = value("-a (--algorithm)");
      if (a.algo != "gemm" && a.algo != "fft" && a.algo != "fwht")
        throw ArgError("Value '" + a.algo + "' does not meet constraint: gemm|fft|fwht",
                       "-a (--algorithm)");
    } else if (s == "--seed") {
      const std::string v = value("--seed");
      try {
        size_t used = 0;
        a.seed = std::stoull(v, &used);
        if (used != v.size() || v[0] == '-') throw std::invalid_argument(v);
      } catch (const std::exception&) {
        throw ArgError("Couldn't read argument value from string '" + v + "'", "--seed");
      }
      a.seed_set = true;
    } else if (s == "--hash") {
      const std::string v = value("--hash");
      if (v == "multiply add shift") a.family = CM_MULTIPLY_ADD_SHIFT;
      else if (v == "multiply shift") a.family = CM_MULTIPLY_SHIFT;
      else if (v == "tabulation8") a.family = CM_TABULATION8;
      else if (v == "tabulation16") a.family = CM_TABULATION16;
      else throw ArgError("Invalid hash `" + v + "'", "--hash");
    } else if (s == "-f" || s == "--force") {
      a.force = true;
    } else if (s == "-v" || s == "--verbose") {
      ++a.verbose;
    } else if (s.size() > 2 && s[0] == '-' && s[1] != '-' &&
               s.find_first_not_of("fv", 1) == std::string::npos) {
      for (size_t k = 1; k < s.size(); ++k) {  // combined short switches, e.g. -fvv
        if (s[k] == 'f')
          a.force = true;
        else
          ++a.verbose;
      }
    } else {
      throw ArgError("Couldn't find match for argument", s);
    }
  }
  static const char* names[3] = {"<A.npy>", "<B.npy>", "<C.npy>"};
  if (a.files.size() < 3) throw ArgError("Required argument missing: " + std::string(names[a.files.size()]), "undefined");
  if (a.files.size() > 3) throw ArgError("Couldn't find match for argument", a.files[3]);
  return a;
}

cmnpy::Array load_operand(const std::string& path, const char* name) {
  cmnpy::Array x = cmnpy::load(path);
  if (x.shape.size() != 2)
    throw std::runtime_error(std::string("Dimension mismatch: ") + name + " has " +
                             std::to_string(x.shape.size()) + " dims, but 2 expected!");
  if (x.word_size != sizeof(double))
    throw std::runtime_error(std::string("Word size mismatch: ") + name +
                             " is expected to have word size of " + std::to_string(sizeof(double)) +
                             ", but got " + std::to_string(x.word_size));
  if (x.descr != "<f8" && x.descr != "=f8")
    throw std::runtime_error(std::string("Type mismatch: ") + name +
                             " is expected to hold little-endian float64, but has descr " + x.descr);
  return x;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Args args = parse(argc, argv);
    const std::string& aFile = args.files[0];
    const std::string& bFile = args.files[1];
    const std::string& cFile = args.files[2];
    if (std::filesystem::exists(std::filesystem::path(cFile)) && !args.force) {
      std::cerr << argv[0] << ": error: file `" << cFile << "' already exists!" << std::endl;
      return EXIT_FAILURE;
    }
    const int verbose = args.verbose;
    const int bParam = args.b, dParam = args.d;
    if (args.algo != "gemm") {
      if (!args.b_set)
        throw std::runtime_error("Got algorithm `" + args.algo + "' but the parameter b is not set!");
      if (bParam < 1 || (bParam & (bParam - 1)) != 0)
        throw std::runtime_error("The parameter b must be a positive power of 2 (got " +
                                 std::to_string(bParam) + ")");
      if (!args.d_set)
        throw std::runtime_error("Got algorithm `" + args.algo + "' but the parameter d is not set!");
      if (dParam < 1)
        throw std::runtime_error("The parameter d must be positive (got " + std::to_string(dParam) + ")");
    }

    cmnpy::Array a = load_operand(aFile, "A");
    const int q = int(a.shape[0]), p = int(a.shape[1]);
    if (a.fortran_order) {
      if (verbose > 0) std::cerr << "A was given in Fortran order, converting to C order" << std::endl;
      fortran_to_c(q, p, a.as<double>());
    }
    cmnpy::Array b = load_operand(bFile, "B");
    if (int(b.shape[0]) != q)
      throw std::runtime_error("Dimension mismatch: B is expected to have " + std::to_string(q) +
                               " rows, but got " + std::to_string(b.shape[0]) + " instead");
    const int r = int(b.shape[1]);
    if (b.fortran_order) {
      if (verbose > 0) std::cerr << "B was given in Fortran order, converting to C order" << std::endl;
      fortran_to_c(q, r, b.as<double>());
    }

    if (verbose > 0) {
      std::cerr << "A = " << aFile << std::endl
                << "B = " << bFile << std::endl
                << "C = " << cFile << std::endl
                << "Force overwrite? " << (args.force ? "Yes" : "No") << std::endl
                << "Algorithm: " << args.algo << std::endl
                << "b = " << bParam << std::endl
                << "d = " << dParam << std::endl;
      std::cerr << aFile << ": " << p << "x" << q << " (" << (a.fortran_order ? "Fortran" : "C")
                << " order, word size = " << a.word_size << ")" << std::endl;
      if (verbose > 1) {
        std::cerr << "A:" << std::endl;
        print_matrix_t(stderr, a.as<double>(), q, p);
      }
      std::cerr << bFile << ": " << q << "x" << r << " (" << (b.fortran_order ? "Fortran" : "C")
                << " order, word size = " << b.word_size << ")" << std::endl;
      if (verbose > 1) {
        std::cerr << "B:" << std::endl;
        print_matrix(stderr, b.as<double>(), q, r);
      }
    }

    std::vector<double> c(size_t(p) * r);
    int rc;
    if (args.algo == "gemm") {
      if (verbose > 0) std::cerr << "Using dgemm multiplication" << std::endl;
      rc = cm_gemm(p, q, r, c.data(), a.as<double>(), b.as<double>());
    } else if (args.algo == "fft") {
      if (verbose > 0) std::cerr << "Using compressed FFT multiplication" << std::endl;
      rc = cm_matmul_fft(p, q, r, c.data(), a.as<double>(), b.as<double>(), dParam, bParam,
                         args.seed, args.seed_set, args.family);
    } else {
      if (verbose > 0) std::cerr << "Using compressed FWHT multiplication" << std::endl;
      rc = cm_matmul_fwht(p, q, r, c.data(), a.as<double>(), b.as<double>(), dParam, bParam,
                          args.seed, args.seed_set, args.family);
    }
    if (rc != CM_OK) throw std::runtime_error(cm_last_error());

    if (verbose > 1) {
      std::cerr << "C:" << std::endl;
      print_matrix(stderr, c.data(), p, r);
    }
    if (verbose > 0) std::cerr << "Saving result to " << cFile << std::endl;
    cmnpy::save_f64(cFile, c.data(), {size_t(p), size_t(r)});
  } catch (const ArgError& e) {
    std::cerr << argv[0] << ": error: " << e.what() << " for arg " << e.id << std::endl;
    return EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::cerr << argv[0] << ": error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
