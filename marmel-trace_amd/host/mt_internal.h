/*
 * mt_internal.h - the host scene container shared by the host library's
 * translation units (not part of the ABI).
 */
#ifndef MT_INTERNAL_H
#define MT_INTERNAL_H

#include "marmel_host.h"

struct mt_scene {
    mh_prim *prims;
    mh_prim_attr *attrs;
    int n, cap;
    mh_material *mats;
    int nm, capm;
    mh_sky sky;
    mh_fog fog;
    mh_ocean ocean;
    double water_level;
    int water_material;
    /* BVH */
    mh_bvh_node *nodes;
    int n_nodes;
    int *leaf_prims;
    int n_leaf;
    int *planes;
    int n_planes;
    int bvh_valid;
};

/* Set the thread's error text; returns code. */
int mt_set_err(int code, const char *fmt, ...);

/* Make room for `extra` more prims (with attributes when attrs is set) and
 * return the index of the first new slot; the caller fills
 * prims[first .. first + extra) and then sets s->n. < 0 on failure. */
long mt_reserve(mt_scene *s, long extra, int attrs);

#endif
