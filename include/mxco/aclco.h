// Source-compatibility spelling for code written against the reference's
// headers: `aclco::StaticMap<...>` etc. name the mxco containers.  Include
// this instead of the reference's static_map.h / static_set.h / dynamic_map.h /
// static_multimap.h / static_multiset.h / bloom_filter.h / hyperloglog.h /
// roaring_bitmap.h; the stream arguments become hipStream_t.
#pragma once

#include "bloom_filter.h"
#include "dynamic_map.h"
#include "hyperloglog.h"
#include "roaring_bitmap.h"
#include "static_map.h"
#include "static_multimap.h"
#include "static_multiset.h"
#include "static_set.h"

namespace aclco = mxco;
