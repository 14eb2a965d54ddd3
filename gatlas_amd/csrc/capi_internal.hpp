// Definitions shared by the extern "C" translation units.
#pragma once

#include "../../include/gatlas.h"
#include "context.hpp"

struct gat_ctx {
    gat::Context impl;
};

// Tile configurations across the ABI: wk 0 (a caller that leaves the field
// zero) means no k-split, as 1 does.
inline gat::MfmaCfg from_tile_cfg(const gat_tile_cfg& t) {
    return gat::MfmaCfg{t.waves_m, t.waves_n, t.tm, t.tn, t.bk, t.blk, t.wk > 1 ? t.wk : 1};
}
inline gat_tile_cfg to_tile_cfg(const gat::MfmaCfg& c) {
    return gat_tile_cfg{c.waves_m, c.waves_n, c.tm, c.tn, c.bk, c.blk, c.wk};
}
