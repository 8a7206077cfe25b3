// Standard headers the reference's translation units include (besides
// <Rcpp.h>, which none of the extracted functions use).  build_ref.sh puts
// this in front of the line ranges it compiles into oracle/_ref.
//
// TEST INFRASTRUCTURE ONLY: the product library never links or loads _ref.
#include <algorithm>
#include <array>
#include <atomic>
#include <cctype>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <utility>
#include <vector>
