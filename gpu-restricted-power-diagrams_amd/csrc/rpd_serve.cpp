// rpd_serve.cpp -- placeholder, replaced by the serving loop.
#include "rpd.h"
