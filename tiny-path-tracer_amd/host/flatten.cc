// Flattener implementation (see flatten.h).
#include "flatten.h"

#include <cstring>
#include <stdexcept>

#include "../scene/camera.h"
#include "../scene/hitable.h"
#include "../scene/hitable_list.h"
#include "../scene/material.h"
#include "../scene/perlin_noise.h"
#include "../scene/rect_box.h"
#include "../scene/sphere.h"
#include "../scene/texture.h"

namespace tpt {
namespace {

struct flatten_error : std::runtime_error {
  int code;
  flatten_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float bitsf(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// One wrapper operation of a chain (outermost first).
struct xop {
  uint32_t kind;
  float p[3];
  bool operator<(const xop &o) const {
    if (kind != o.kind) return kind < o.kind;
    return std::memcmp(p, o.p, sizeof p) < 0;
  }
};

// A primitive seen through a chain of wrappers.
struct prim_ref {
  const hitable *prim;
  std::vector<xop> chain;
  uint32_t flip;
};

class flattener {
 public:
  explicit flattener(flat_scene *out) : out_(out) {}

  void world(const hitable *w) {
    if (!w) throw flatten_error(TPT_ERR_INVALID, "world is null");
    std::vector<const hitable *> elems;
    if (dynamic_cast<const hitable_list *>(w) && !dynamic_cast<const bvh_node *>(w)) {
      // Top-level list: nested lists flatten into one sequence (the shrinking
      // t_max threads through them unchanged); bvh_node elements (bare or
      // seen through translate / rotate_y / flip_normal) become their own
      // segments, runs of other elements one LIST leaf segment.
      expand_list(w, &elems);
    } else {
      elems.push_back(w);
    }
    size_t i = 0;
    while (i < elems.size()) {
      const uint32_t start = static_cast<uint32_t>(out_->slots.size());
      std::vector<xop> chain;
      uint32_t flip = 0;
      const hitable *inner = peel(elems[i], &chain, &flip);
      uint32_t word = TPT_SEGMENT_PLAIN;
      if (dynamic_cast<const bvh_node *>(inner)) {
        if (!chain.empty() || flip) {
          if (chain.size() > TPT_XFORM_MAX_OPS)
            throw flatten_error(TPT_ERR_UNSUPPORTED, "more than 4 nested translate/rotate_y wrappers");
          word = (chain.empty() ? TPT_SEGMENT_PLAIN : xform_id(chain)) | (flip ? TPT_SEGMENT_FLIP : 0u);
        }
        node(inner, 1);
        ++i;
      } else if (elems.size() == 1) {
        node(elems[i], 0);
        ++i;
      } else {
        std::vector<prim_ref> refs;
        while (i < elems.size() && !dynamic_cast<const bvh_node *>(peel(elems[i], nullptr, nullptr))) {
          collect(elems[i], {}, 0, &refs);
          ++i;
        }
        emit_leaf(refs);
      }
      out_->segments.push_back(start);
      out_->segments.push_back(static_cast<uint32_t>(out_->slots.size()));
      out_->segment_xforms.push_back(word);
    }
    bool wrapped = false;
    for (uint32_t x : out_->segment_xforms) wrapped = wrapped || x != TPT_SEGMENT_PLAIN;
    if (!wrapped) out_->segment_xforms.clear();
  }

  void lights(const hitable *l) {
    if (!l) return;
    out_->light_mode = 0;
    if (auto *lst = dynamic_cast<const hitable_list *>(l)) {
      out_->light_mode = 1;
      for (int i = 0; i < lst->list_size_; ++i)
        if (dynamic_cast<const hitable_list *>(lst->list_[i])) out_->light_mode = 2;
    }
    light(l, 1.0f);
  }

 private:
  flat_scene *out_;
  bool in_boundary_ = false;     // collecting a medium's boundary
  bool skip_material_ = false;   // boundary primitives number no material
  std::vector<prim_ref> pending_boundary_;  // set by prim_words for a medium
  std::map<std::vector<xop>, uint32_t> xform_ids_;
  std::map<const material *, uint32_t> material_ids_;
  std::map<const texture *, uint32_t> texture_ids_;

  // The hitable under a chain of translate / rotate_y / flip_normal wrappers
  // (chain outermost first; null outputs are not filled).
  static const hitable *peel(const hitable *h, std::vector<xop> *chain, uint32_t *flip) {
    for (;;) {
      if (auto *f = dynamic_cast<const flip_normal *>(h)) {
        if (flip) *flip ^= 1u;
        h = f->ptr_;
      } else if (auto *t = dynamic_cast<const translate *>(h)) {
        if (chain) chain->push_back(xop{TPT_XFORM_TRANSLATE, {t->offset_.e[0], t->offset_.e[1], t->offset_.e[2]}});
        h = t->ptr_;
      } else if (auto *r = dynamic_cast<const rotate_y *>(h)) {
        if (chain) chain->push_back(xop{TPT_XFORM_ROTATE_Y, {r->sin_theta_, r->cos_theta_, 0.0f}});
        h = r->ptr_;
      } else {
        return h;
      }
    }
  }

  static void expand_list(const hitable *h, std::vector<const hitable *> *out) {
    const auto *lst = dynamic_cast<const hitable_list *>(h);
    if (lst) {
      for (int i = 0; i < lst->list_size_; ++i) expand_list(lst->list_[i], out);
    } else {
      out->push_back(h);
    }
  }

  // Pre-order emission of a bvh_node tree (depth counts bvh_node levels).
  void node(const hitable *h, uint32_t depth) {
    const auto *bn = dynamic_cast<const bvh_node *>(h);
    if (!bn) {
      std::vector<prim_ref> refs;
      collect(h, {}, 0, &refs);
      emit_leaf(refs);
      return;
    }
    if (depth > out_->max_depth) out_->max_depth = depth;
    if (!bn->left_ || !bn->right_) throw flatten_error(TPT_ERR_INVALID, "bvh_node with a null child");
    const uint32_t idx = static_cast<uint32_t>(out_->slots.size());
    tpt_slot s{};
    for (int i = 0; i < 3; ++i) {
      s.f[i] = bn->box_.min_.e[i];
      s.f[3 + i] = bn->box_.max_.e[i];
    }
    out_->slots.push_back(s);
    if (bn->right_ == bn->left_ && dynamic_cast<const constant_medium *>(bn->left_))
      // the reference evaluates that child twice, each time with its own draw
      throw flatten_error(TPT_ERR_UNSUPPORTED, "constant_medium as the only child of a bvh_node");
    child(bn->left_, depth);
    uint32_t right = TPT_NO_RIGHT;
    if (bn->right_ != bn->left_) {  // n == 1 stores one child twice; it is evaluated once
      right = static_cast<uint32_t>(out_->slots.size());
      child(bn->right_, depth);
    }
    out_->slots[idx].a = static_cast<uint32_t>(out_->slots.size());
    out_->slots[idx].b = (TPT_SLOT_BOX << 28) | right;
  }

  void child(const hitable *h, uint32_t depth) {
    if (dynamic_cast<const bvh_node *>(h)) {
      node(h, depth + 1);
    } else {
      std::vector<prim_ref> refs;
      collect(h, {}, 0, &refs);
      emit_leaf(refs);
    }
  }

  // Gathers the primitives under a non-BVH hitable in hit() evaluation order.
  void collect(const hitable *h, std::vector<xop> chain, uint32_t flip, std::vector<prim_ref> *refs) {
    if (!h) throw flatten_error(TPT_ERR_INVALID, "null hitable in scene graph");
    if (auto *f = dynamic_cast<const flip_normal *>(h)) {
      collect(f->ptr_, chain, flip ^ 1u, refs);
    } else if (auto *t = dynamic_cast<const translate *>(h)) {
      chain.push_back(xop{TPT_XFORM_TRANSLATE, {t->offset_.e[0], t->offset_.e[1], t->offset_.e[2]}});
      collect(t->ptr_, chain, flip, refs);
    } else if (auto *r = dynamic_cast<const rotate_y *>(h)) {
      chain.push_back(xop{TPT_XFORM_ROTATE_Y, {r->sin_theta_, r->cos_theta_, 0.0f}});
      collect(r->ptr_, chain, flip, refs);
    } else if (auto *b = dynamic_cast<const box *>(h)) {
      collect(b->list_ptr_, chain, flip, refs);
    } else if (dynamic_cast<const bvh_node *>(h)) {
      throw flatten_error(TPT_ERR_UNSUPPORTED,
                          "bvh_node below a wrapper or list inside a bvh_node is not supported");
    } else if (auto *l = dynamic_cast<const hitable_list *>(h)) {
      for (int i = 0; i < l->list_size_; ++i) collect(l->list_[i], chain, flip, refs);
    } else if (dynamic_cast<const sphere *>(h) || dynamic_cast<const xy_rect *>(h) ||
               dynamic_cast<const xz_rect *>(h) || dynamic_cast<const yz_rect *>(h) ||
               dynamic_cast<const moving_sphere *>(h)) {
      if (chain.size() > TPT_XFORM_MAX_OPS)
        throw flatten_error(TPT_ERR_UNSUPPORTED, "more than 4 nested translate/rotate_y wrappers");
      refs->push_back(prim_ref{h, chain, flip});
    } else if (dynamic_cast<const constant_medium *>(h)) {
      if (!chain.empty() || flip)
        throw flatten_error(TPT_ERR_UNSUPPORTED, "constant_medium below translate/rotate_y/flip_normal");
      if (in_boundary_) throw flatten_error(TPT_ERR_UNSUPPORTED, "constant_medium inside a medium's boundary");
      refs->push_back(prim_ref{h, chain, flip});
    } else {
      throw flatten_error(TPT_ERR_UNSUPPORTED, "unknown hitable type in scene graph");
    }
  }

  // Primitive parameters + kind/axis/material word shared by slots and entries.
  void prim_words(const hitable *h, uint32_t flip, float f[5], uint32_t *a, uint32_t *kind) {
    uint32_t mat, axis = 0;
    if (auto *s = dynamic_cast<const sphere *>(h)) {
      f[0] = s->center_.e[0];
      f[1] = s->center_.e[1];
      f[2] = s->center_.e[2];
      f[3] = s->radius_;
      f[4] = 0.0f;
      mat = material_id(s->mat_ptr_);
      *kind = TPT_SLOT_SPHERE;
    } else if (auto *cm = dynamic_cast<const constant_medium *>(h)) {
      // boundary primitives are appended by emit_leaf; f[0] (first entry) is
      // patched there, f[1] = count
      mat = material_id(cm->phase_funcion_);
      in_boundary_ = true;
      std::vector<prim_ref> b;
      collect(cm->boundary_, {}, 0, &b);
      in_boundary_ = false;
      if (b.empty()) throw flatten_error(TPT_ERR_INVALID, "constant_medium with an empty boundary");
      f[0] = 0.0f;
      f[1] = bitsf(static_cast<uint32_t>(b.size()));
      f[2] = cm->density_;
      f[3] = f[4] = 0.0f;
      pending_boundary_ = std::move(b);
      *kind = TPT_SLOT_MEDIUM;
    } else if (auto *ms = dynamic_cast<const moving_sphere *>(h)) {
      tpt_motion m{};
      for (int i = 0; i < 3; ++i) m.center0[i] = ms->center0_.e[i], m.center1[i] = ms->center1_.e[i];
      m.radius = ms->radius_;
      m.time0 = ms->time0_;
      m.time1 = ms->time1_;
      f[0] = bitsf(static_cast<uint32_t>(out_->motions.size()));
      f[1] = f[2] = f[3] = f[4] = 0.0f;
      out_->motions.push_back(m);
      mat = material_id(ms->mat_ptr_);
      *kind = TPT_SLOT_MSPHERE;
    } else {
      *kind = TPT_SLOT_RECT;
      if (auto *r = dynamic_cast<const xy_rect *>(h)) {
        axis = 2;
        f[0] = r->k_, f[1] = r->x0_, f[2] = r->x1_, f[3] = r->y0_, f[4] = r->y1_;
        mat = material_id(r->mat_ptr_);
      } else if (auto *r2 = dynamic_cast<const xz_rect *>(h)) {
        axis = 1;
        f[0] = r2->k_, f[1] = r2->x0_, f[2] = r2->x1_, f[3] = r2->z0_, f[4] = r2->z1_;
        mat = material_id(r2->mat_ptr_);
      } else {
        auto *r3 = dynamic_cast<const yz_rect *>(h);
        axis = 0;
        f[0] = r3->k_, f[1] = r3->y0_, f[2] = r3->y1_, f[3] = r3->z0_, f[4] = r3->z1_;
        mat = material_id(r3->mat_ptr_);
      }
    }
    *a = mat | (axis << 28) | (flip << 30);
  }

  // A medium boundary's primitives: only their geometry is used (the medium
  // hit takes its material from the phase function), so no material is
  // numbered for them.
  void boundary_entries(const std::vector<prim_ref> &refs) {
    for (const prim_ref &r : refs) {
      tpt_entry e{};
      uint32_t kind, a;
      skip_material_ = true;
      prim_words(r.prim, r.flip, e.f, &a, &kind);
      skip_material_ = false;
      e.a = a;
      e.b = kind;
      e.xform = r.chain.empty() ? 0xFFFFFFFFu : xform_id(r.chain);
      out_->entries.push_back(e);
    }
  }

  void emit_leaf(const std::vector<prim_ref> &refs) {
    const uint32_t leaf = out_->n_leaves;
    if (leaf + refs.size() > TPT_PAYLOAD_MASK) throw flatten_error(TPT_ERR_UNSUPPORTED, "too many leaves");
    tpt_slot s{};
    if (refs.size() == 1 && refs[0].chain.empty()) {
      float f[5];
      uint32_t a, kind;
      prim_words(refs[0].prim, refs[0].flip, f, &a, &kind);
      if (kind == TPT_SLOT_MEDIUM) {
        f[0] = bitsf(static_cast<uint32_t>(out_->entries.size()));
        boundary_entries(pending_boundary_);
      }
      for (int i = 0; i < 5; ++i) s.f[i] = f[i];
      s.a = a;
      s.b = (kind << 28) | leaf;
      out_->n_leaves += 1;
    } else {
      s.a = static_cast<uint32_t>(out_->entries.size());
      s.f[0] = bitsf(leaf);
      s.f[1] = bitsf(static_cast<uint32_t>(refs.size()));
      s.b = TPT_SLOT_LIST << 28;
      // the list's entries are contiguous; media boundaries follow them
      std::vector<std::pair<size_t, std::vector<prim_ref>>> media;
      for (const prim_ref &r : refs) {
        tpt_entry e{};
        uint32_t kind;
        prim_words(r.prim, r.flip, e.f, &e.a, &kind);
        e.b = kind;
        e.xform = r.chain.empty() ? 0xFFFFFFFFu : xform_id(r.chain);
        if (kind == TPT_SLOT_MEDIUM) media.emplace_back(out_->entries.size(), std::move(pending_boundary_));
        out_->entries.push_back(e);
      }
      for (auto &m : media) {
        out_->entries[m.first].f[0] = bitsf(static_cast<uint32_t>(out_->entries.size()));
        boundary_entries(m.second);
      }
      out_->n_leaves += static_cast<uint32_t>(refs.size());
    }
    out_->slots.push_back(s);
  }

  uint32_t xform_id(const std::vector<xop> &chain) {
    auto it = xform_ids_.find(chain);
    if (it != xform_ids_.end()) return it->second;
    tpt_xform x{};
    x.nops = static_cast<uint32_t>(chain.size());
    for (size_t i = 0; i < chain.size(); ++i) {
      x.kinds |= chain[i].kind << (2 * i);
      for (int j = 0; j < 3; ++j) x.p[i][j] = chain[i].p[j];
    }
    const uint32_t id = static_cast<uint32_t>(out_->xforms.size());
    out_->xforms.push_back(x);
    xform_ids_[chain] = id;
    return id;
  }

  uint32_t material_id(const material *m) {
    if (skip_material_) return 0;
    if (!m) throw flatten_error(TPT_ERR_INVALID, "scene primitive with a null material");
    auto it = material_ids_.find(m);
    if (it != material_ids_.end()) return it->second;
    tpt_material rec{};
    if (auto *l = dynamic_cast<const lambertian *>(m)) {
      rec.kind = TPT_MAT_LAMBERTIAN;
      rec.tex = texture_id(l->albedo_);
    } else if (auto *me = dynamic_cast<const metal *>(m)) {
      rec.kind = TPT_MAT_METAL;
      for (int i = 0; i < 3; ++i) rec.p[i] = me->albedo_.e[i];
      rec.p[3] = me->fuzz_;
    } else if (auto *d = dynamic_cast<const dielectric *>(m)) {
      rec.kind = TPT_MAT_DIELECTRIC;
      rec.p[0] = d->ref_idx_;
    } else if (auto *dl = dynamic_cast<const diffuse_light *>(m)) {
      rec.kind = TPT_MAT_DIFFUSE_LIGHT;
      rec.tex = texture_id(dl->emit_);
    } else if (auto *iso = dynamic_cast<const isotropic *>(m)) {
      rec.kind = TPT_MAT_ISOTROPIC;
      rec.tex = texture_id(iso->albedo_);
    } else {
      throw flatten_error(TPT_ERR_UNSUPPORTED, "material type not on the device path");
    }
    const uint32_t id = static_cast<uint32_t>(out_->materials.size());
    if (id >= (1u << 28)) throw flatten_error(TPT_ERR_UNSUPPORTED, "too many materials");
    out_->materials.push_back(rec);
    material_ids_[m] = id;
    return id;
  }

  uint32_t texture_id(const texture *t) {
    if (!t) throw flatten_error(TPT_ERR_INVALID, "null texture");
    auto it = texture_ids_.find(t);
    if (it != texture_ids_.end()) return it->second;
    const uint32_t id = static_cast<uint32_t>(out_->textures.size());
    out_->textures.push_back(tpt_texture{});
    texture_ids_[t] = id;
    tpt_texture rec{};
    if (auto *c = dynamic_cast<const constant_texture *>(t)) {
      rec.kind = TPT_TEX_CONSTANT;
      for (int i = 0; i < 3; ++i) rec.p[i] = c->color_.e[i];
    } else if (auto *ch = dynamic_cast<const checker_texture *>(t)) {
      rec.kind = TPT_TEX_CHECKER;
      rec.a = texture_id(ch->odd_);
      rec.b = texture_id(ch->even_);
    } else if (auto *pn = dynamic_cast<const perlin_noise_texture *>(t)) {
      rec.kind = TPT_TEX_PERLIN;
      rec.p[0] = pn->scale_;
      rec.a = static_cast<uint32_t>(pn->octaves_ < 0 ? 0 : pn->octaves_);
      snapshot_perlin();
    } else if (auto *im = dynamic_cast<const image_texture *>(t)) {
      rec.kind = TPT_TEX_IMAGE;
      if (!im->data_ || im->width_ <= 0 || im->height_ <= 0)
        throw flatten_error(TPT_ERR_INVALID, "image_texture without pixels");
      rec.a = static_cast<uint32_t>(out_->texels.size());
      rec.b = static_cast<uint32_t>(im->width_);
      rec.c = static_cast<uint32_t>(im->height_);
      const size_t n = static_cast<size_t>(im->width_) * im->height_ * 3;
      if (out_->texels.size() + n > 0xFFFFFFFFull) throw flatten_error(TPT_ERR_UNSUPPORTED, "texel pool over 4 GiB");
      out_->texels.insert(out_->texels.end(), im->data_, im->data_ + n);
    } else {
      throw flatten_error(TPT_ERR_UNSUPPORTED, "texture type not on the device path");
    }
    out_->textures[id] = rec;
    return id;
  }

  void snapshot_perlin() {
    if (!out_->perlin.empty()) return;
    tpt_perlin p{};
    for (int i = 0; i < 256; ++i) {
      for (int j = 0; j < 3; ++j) p.grad[i][j] = perlin_noise::random_vec3_[i].e[j];
      p.perm[0][i] = perlin_noise::permute_x_[i];
      p.perm[1][i] = perlin_noise::permute_y_[i];
      p.perm[2][i] = perlin_noise::permute_z_[i];
    }
    out_->perlin.push_back(p);
  }

  // Light list: hitable_list pdf = mean of its children (src/hitable_list.cc:24-36);
  // xz_rect and sphere have pdfs; any other hitable keeps the base class
  // defaults (pdf 0, direction (1,0,0); headers/hitable.h:35-38).
  void light(const hitable *h, float weight) {
    tpt_light rec{};
    rec.weight = weight;
    if (auto *l = dynamic_cast<const hitable_list *>(h)) {
      const float w = static_cast<float>(1.0 / l->list_size_);
      for (int i = 0; i < l->list_size_; ++i) light(l->list_[i], weight * w);
      return;
    }
    if (auto *r = dynamic_cast<const xz_rect *>(h)) {
      rec.kind = TPT_LIGHT_XZ_RECT;
      rec.p[0] = r->x0_, rec.p[1] = r->x1_, rec.p[2] = r->z0_, rec.p[3] = r->z1_, rec.p[4] = r->k_;
    } else if (auto *s = dynamic_cast<const sphere *>(h)) {
      rec.kind = TPT_LIGHT_SPHERE;
      for (int i = 0; i < 3; ++i) rec.p[i] = s->center_.e[i];
      rec.p[3] = s->radius_;
    } else {
      rec.kind = TPT_LIGHT_NONE;
    }
    out_->lights.push_back(rec);
  }
};

}  // namespace

tpt_flat_scene flat_scene::view() const {
  tpt_flat_scene v{};
  v.slots = slots.data();
  v.n_slots = static_cast<uint32_t>(slots.size());
  v.segments = segments.data();
  v.n_segments = static_cast<uint32_t>(segments.size() / 2);
  v.entries = entries.data();
  v.n_entries = static_cast<uint32_t>(entries.size());
  v.xforms = xforms.data();
  v.n_xforms = static_cast<uint32_t>(xforms.size());
  v.materials = materials.data();
  v.n_materials = static_cast<uint32_t>(materials.size());
  v.textures = textures.data();
  v.n_textures = static_cast<uint32_t>(textures.size());
  v.texels = texels.data();
  v.n_texels = texels.size();
  v.perlin = perlin.empty() ? nullptr : perlin.data();
  v.lights = lights.data();
  v.n_lights = static_cast<uint32_t>(lights.size());
  v.light_mode = light_mode;
  v.n_leaves = n_leaves;
  v.max_depth = max_depth;
  for (int i = 0; i < 6; ++i) v.bounds[i] = bounds[i];
  v.motions = motions.data();
  v.n_motions = static_cast<uint32_t>(motions.size());
  v.segment_xforms = segment_xforms.empty() ? nullptr : segment_xforms.data();
  return v;
}

int flatten(const hitable *world, const hitable *lights, flat_scene *out, std::string *err) {
  *out = flat_scene{};
  try {
    flattener f(out);
    f.world(world);
    f.lights(lights);
    AABB box;
    if (world->bounding_box(0, 1, box)) {
      for (int i = 0; i < 3; ++i) {
        out->bounds[i] = box.min_.e[i];
        out->bounds[3 + i] = box.max_.e[i];
      }
    }
    return TPT_OK;
  } catch (const flatten_error &e) {
    if (err) *err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    if (err) *err = e.what();
    return TPT_ERR_INVALID;
  }
}

tpt_camera camera_fields(const camera_with_blur &cam) {
  tpt_camera c{};
  for (int i = 0; i < 3; ++i) {
    c.origin[i] = cam.origin_.e[i];
    c.lower_left[i] = cam.lower_left_corner_.e[i];
    c.horizontal[i] = cam.horizontal_.e[i];
    c.vertical[i] = cam.vertical_.e[i];
    c.u[i] = cam.u_.e[i];
    c.v[i] = cam.v_.e[i];
    c.w[i] = cam.w_.e[i];
  }
  c.lens_radius = cam.lens_radius_;
  c.time0 = cam.time0;
  c.time1 = cam.time1;
  return c;
}

}  // namespace tpt
