#include "index.hpp"
