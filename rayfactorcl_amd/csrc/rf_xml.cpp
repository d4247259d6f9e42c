// rf_xml.cpp -- host-side scene input and result output of the C ABI.
//
// A small XML reader (elements, attributes, comments, declarations, text)
// and the scene semantics of RayFactorCL/Scene.cpp:24-452 with the affine
// builder of "Affine Transformation/AffineTransformation.cpp:14-111" and
// "Matrices.cpp:22-58".  The per-type rayDensity expressions, the float
// operation order of the matrix products and the libm calls are kept as the
// reference evaluates them, so the Primitive / BoundingVolume bytes are
// identical (fields the reference leaves uninitialised are zero here).
// Malformed input returns RF_ERR_PARSE instead of crashing or exiting.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "rayfactor.h"

namespace {

// ---------------------------------------------------------------------------
// Minimal DOM
// ---------------------------------------------------------------------------
struct Node {
    enum Kind { ELEMENT, TEXT, COMMENT, DECL, OTHER } kind = ELEMENT;
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<Node>> children;

    const char *attr(const char *key) const
    {
        for (const auto &a : attrs)
            if (a.first == key) return a.second.c_str();
        return nullptr;
    }
    const Node *first_child() const { return children.empty() ? nullptr : children[0].get(); }
    const Node *first_child_element(const char *nm = nullptr) const
    {
        for (const auto &c : children)
            if (c->kind == ELEMENT && (!nm || c->name == nm)) return c.get();
        return nullptr;
    }
};

class Parser {
public:
    explicit Parser(const std::string &s) : s_(s), p_(0) {}

    bool parse(Node &doc)
    {
        doc.kind = Node::ELEMENT;
        return parse_content(doc, nullptr) && ok_;
    }

private:
    const std::string &s_;
    size_t p_;
    bool ok_ = true;

    bool eof() const { return p_ >= s_.size(); }
    static bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
    void skip_ws() { while (!eof() && ws(s_[p_])) ++p_; }
    bool starts(const char *t) const { return s_.compare(p_, std::strlen(t), t) == 0; }

    static std::string decode(const std::string &v)
    {
        std::string o;
        for (size_t i = 0; i < v.size(); ++i) {
            if (v[i] == '&') {
                static const struct { const char *e; char c; } ents[] = {
                    {"&amp;", '&'}, {"&lt;", '<'}, {"&gt;", '>'}, {"&quot;", '"'}, {"&apos;", '\''}};
                bool hit = false;
                for (const auto &en : ents) {
                    const size_t L = std::strlen(en.e);
                    if (v.compare(i, L, en.e) == 0) { o += en.c; i += L - 1; hit = true; break; }
                }
                if (!hit) o += v[i];
            } else {
                o += v[i];
            }
        }
        return o;
    }

    bool parse_name(std::string &out)
    {
        const size_t b = p_;
        while (!eof() && !ws(s_[p_]) && s_[p_] != '>' && s_[p_] != '/' && s_[p_] != '=') ++p_;
        out = s_.substr(b, p_ - b);
        return !out.empty();
    }

    // Parses children until the closing tag of `open` (or end of input at top level).
    bool parse_content(Node &parent, const std::string *open)
    {
        for (;;) {
            skip_ws();
            if (eof()) return open == nullptr;
            if (starts("</")) {
                if (!open) return false;
                p_ += 2;
                std::string nm;
                if (!parse_name(nm) || nm != *open) return false;
                skip_ws();
                if (eof() || s_[p_] != '>') return false;
                ++p_;
                return true;
            }
            auto node = std::make_unique<Node>();
            if (starts("<?")) {
                const size_t e = s_.find("?>", p_);
                if (e == std::string::npos) return false;
                node->kind = Node::DECL;
                p_ = e + 2;
            } else if (starts("<!--")) {
                const size_t e = s_.find("-->", p_);
                if (e == std::string::npos) return false;
                node->kind = Node::COMMENT;
                p_ = e + 3;
            } else if (starts("<![CDATA[")) {
                const size_t e = s_.find("]]>", p_);
                if (e == std::string::npos) return false;
                node->kind = Node::TEXT;
                p_ = e + 3;
            } else if (starts("<!")) {
                const size_t e = s_.find('>', p_);
                if (e == std::string::npos) return false;
                node->kind = Node::OTHER;
                p_ = e + 1;
            } else if (s_[p_] == '<') {
                ++p_;
                node->kind = Node::ELEMENT;
                if (!parse_name(node->name)) return false;
                bool closed = false;
                for (;;) {
                    skip_ws();
                    if (eof()) return false;
                    if (starts("/>")) { p_ += 2; closed = true; break; }
                    if (s_[p_] == '>') { ++p_; break; }
                    std::string key;
                    if (!parse_name(key)) return false;
                    skip_ws();
                    if (eof() || s_[p_] != '=') return false;
                    ++p_;
                    skip_ws();
                    if (eof() || (s_[p_] != '"' && s_[p_] != '\'')) return false;
                    const char q = s_[p_++];
                    const size_t e = s_.find(q, p_);
                    if (e == std::string::npos) return false;
                    node->attrs.emplace_back(key, decode(s_.substr(p_, e - p_)));
                    p_ = e + 1;
                }
                if (!closed && !parse_content(*node, &node->name)) return false;
            } else {
                // character data up to the next markup
                const size_t e = s_.find('<', p_);
                node->kind = Node::TEXT;
                p_ = (e == std::string::npos) ? s_.size() : e;
            }
            parent.children.push_back(std::move(node));
        }
    }
};

// tinyxml2-style attribute reads (FloatAttribute / QueryFloatAttribute /
// BoolAttribute): sscanf-like parsing, value left unchanged on failure.
bool query_float(const Node *e, const char *key, float *v)
{
    const char *a = e->attr(key);
    if (!a) return false;
    float f;
    if (std::sscanf(a, "%f", &f) != 1) return false;
    *v = f;
    return true;
}

float float_attr(const Node *e, const char *key)
{
    float f = 0;
    query_float(e, key, &f);
    return f;
}

bool bool_attr(const Node *e, const char *key)
{
    const char *a = e->attr(key);
    if (!a) return false;
    int iv = -1;
    std::sscanf(a, "%d", &iv);
    if (iv > 0 || std::strcmp(a, "true") == 0) return true;
    return false;
}

// Iteration used by the reference: FirstChild()->ToElement(), then
// NextSiblingElement().  Stops at once when the first child is not an element.
template <typename F>
bool for_each_element(const Node *parent, F &&fn)
{
    const Node *first = parent->first_child();
    if (!first || first->kind != Node::ELEMENT) return true;
    bool started = false;
    for (const auto &c : parent->children) {
        if (c.get() == first) started = true;
        if (!started || c->kind != Node::ELEMENT) continue;
        if (!fn(c.get())) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// 4x4 float matrices (Matrix4.h:12-22, Matrices.cpp:22-58)
// ---------------------------------------------------------------------------
struct Mat4 {
    float m[16];
    Mat4()
    {
        for (int i = 0; i < 16; ++i) m[i] = 0.0f;
        m[0] = m[5] = m[10] = m[15] = 1.0f;
    }
};

// x <- x * a in row-major reading: sum over k of a[4k+r] * x[4c+k] into x[4c+r]
void pre_multiply(Mat4 &x, const Mat4 &a)
{
    const Mat4 tmp = x;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float sum = 0;
            for (int k = 0; k < 4; ++k) sum += a.m[4 * k + r] * tmp.m[4 * c + k];
            x.m[4 * c + r] = sum;
        }
}

// x <- a * x: sum over k of x[4k+r] * a[4c+k] into x[4c+r]
void post_multiply(Mat4 &x, const Mat4 &a)
{
    const Mat4 tmp = x;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float sum = 0;
            for (int k = 0; k < 4; ++k) sum += tmp.m[4 * k + r] * a.m[4 * c + k];
            x.m[4 * c + r] = sum;
        }
}

const float kPi = 3.141592653589793f;   // AffineTransformation.h:19

// AffineTransformation.cpp:14-35 (axis assumed normalised)
void rotate(Mat4 &m, float angle, float x, float y, float z, bool inverse)
{
    Mat4 rm;
    // the reference's unqualified cos/sin resolve to the C library's double
    // functions; inverse (:38-58) flips the sine terms
    const float ang = angle * kPi / 180.0f;
    const float c = (float)::cos((double)ang);
    const float s = (float)::sin((double)ang);
    const float mc = 1.0f - c;
    rm.m[0] = c + mc * x * x;
    rm.m[1] = inverse ? mc * x * y - s * z : mc * x * y + s * z;
    rm.m[2] = inverse ? mc * x * z + s * y : mc * x * z - s * y;
    rm.m[4] = inverse ? mc * y * x + s * z : mc * y * x - s * z;
    rm.m[5] = c + mc * y * y;
    rm.m[6] = inverse ? mc * y * z - s * x : mc * y * z + s * x;
    rm.m[8] = inverse ? mc * z * x - s * y : mc * z * x + s * y;
    rm.m[9] = inverse ? mc * z * y + s * x : mc * z * y - s * x;
    rm.m[10] = c + mc * z * z;
    if (inverse) pre_multiply(m, rm);
    else post_multiply(m, rm);
}

void scale(Mat4 &m, float sx, float sy, float sz, bool inverse)
{
    Mat4 a;
    if (inverse) { a.m[0] = 1.0f / sx; a.m[5] = 1.0f / sy; a.m[10] = 1.0f / sz; pre_multiply(m, a); }
    else { a.m[0] = sx; a.m[5] = sy; a.m[10] = sz; post_multiply(m, a); }
}

void translate(Mat4 &m, float x, float y, float z, bool inverse)
{
    Mat4 a;
    if (inverse) { a.m[3] = -x; a.m[7] = -y; a.m[11] = -z; pre_multiply(m, a); }
    else { a.m[3] = x; a.m[7] = y; a.m[11] = z; post_multiply(m, a); }
}

// ---------------------------------------------------------------------------
// Scene semantics (Scene.cpp)
// ---------------------------------------------------------------------------
const float kRefPi = 3.141592f;   // RayfactorConstants.h:12

struct SceneBuild {
    std::vector<rf_primitive> prims;
    std::vector<rf_bounding_volume> bvs;
    std::string err;
};

// initPrimitive (Scene.cpp:296-390)
bool init_primitive(const Node *e, rf_primitive &o)
{
    if (!e->first_child()) return true;   // no children: transform and rayDensity stay zero
    Mat4 trm, inv;
    float rayDensity = 1;
    o.rayDensity = 1;
    float xScale = 1.0f, yScale = 1.0f, zScale = 1.0f;
    for_each_element(e, [&](const Node *c) {
        const std::string &t = c->name;
        if (t == "rayDensity") {
            rayDensity = float_attr(c, "value");
        } else if (t == "translation") {
            const float x = float_attr(c, "x"), y = float_attr(c, "y"), z = float_attr(c, "z");
            translate(trm, x, y, z, false);
            translate(inv, x, y, z, true);
        } else if (t == "scale") {
            xScale = float_attr(c, "x"); yScale = float_attr(c, "y"); zScale = float_attr(c, "z");
            scale(trm, xScale, yScale, zScale, false);
            scale(inv, xScale, yScale, zScale, true);
        } else if (t == "rotate") {
            const float d = float_attr(c, "degrees");
            const float x = float_attr(c, "x"), y = float_attr(c, "y"), z = float_attr(c, "z");
            rotate(trm, d, x, y, z, false);
            rotate(inv, d, x, y, z, true);
        }
        return true;
    });
    for (int i = 0; i < 12; ++i) { o.m[i] = trm.m[i]; o.invm[i] = inv.m[i]; }
    // per-type ray density, evaluated with the reference's types (Scene.cpp:361-387):
    // its unqualified pow/sqrt are the C library's double functions
    if (o.type == RF_CYLINDER) {
        o.rayDensity = (int)((int)rayDensity * 2.0 * kRefPi * zScale *
                             ::pow(0.5 * (::pow((double)xScale, 2.0) + ::pow((double)yScale, 2.0)), 0.5));
    } else if (o.type == RF_FRUSTUM) {
        const float sm = o.util1 + 1E-5f;
        o.rayDensity = (int)((int)rayDensity *
                             (::sqrt((double)((xScale * xScale + yScale * yScale) * sm * sm)) * (1.f + sm) * zScale * kRefPi) /
                             (::sqrt(2.0) * sm));
    } else if (o.type == RF_SPHERE) {
        const float p = (float)1.6075;   // Knud Thomsen's ellipsoid area formula
        const double px = ::pow((double)xScale, (double)p);
        const double py = ::pow((double)yScale, (double)p);
        const double pz = ::pow((double)zScale, (double)p);
        o.rayDensity = (int)((int)rayDensity *
                             ((double)(4.0f * kRefPi) * ::pow((px * py + px * pz + py * pz) / 3.0f, (double)(1.0f / p))));
    } else if (o.type == RF_ANNULUS) {
        o.rayDensity = (int)(rayDensity * kRefPi * (o.util2 - o.util1) + 0.5f);
    } else if (o.type == RF_RECTANGLE) {
        o.rayDensity = (int)((int)rayDensity * 2.0f * xScale * 2.0f * yScale);
    } else if (o.type == RF_DISC) {
        o.rayDensity = (int)((int)rayDensity * kRefPi * xScale * yScale);
    }
    return true;
}

// Shape type from the "type" attribute and its radius children
// (Scene.cpp:222-273 for primitives, :117-166 for bounding volumes).
bool classify(const Node *e, float &util1, float &util2, int32_t &type, bool is_bv, float boundMod, std::string &err)
{
    const char *prType = e->attr("type");
    if (!prType) { err = "element <" + e->name + "> has no type attribute"; return false; }
    const std::string t = prType;
    if (t == "cylinderSurface" || t == "taperedCylinderSurface" || t == "frustum") {
        float smallRadius = 1.f;
        const Node *sr = e->first_child_element("smallRadius");
        if (sr) query_float(sr, "value", &smallRadius);
        if (smallRadius == 1.0) {
            type = RF_CYLINDER;
        } else {
            type = RF_FRUSTUM;
            util1 = smallRadius;
        }
        if (!is_bv) util2 = boundMod;
    } else if (t == "annulus") {
        const Node *sr = e->first_child_element("smallRadius");
        const Node *lr = e->first_child_element("largeRadius");
        float smallRadius, largeRadius;
        if (!sr || !lr || !query_float(sr, "value", &smallRadius) || !query_float(lr, "value", &largeRadius)) {
            err = "annulus without smallRadius/largeRadius values";
            return false;
        }
        type = RF_ANNULUS;
        util1 = smallRadius * smallRadius;
        util2 = largeRadius * largeRadius;
    } else if (t == "rectangle") {
        type = RF_RECTANGLE;
    } else if (t == "disc") {
        type = RF_DISC;
    } else if (t == "sphere") {
        type = RF_SPHERE;
        if (!is_bv) util2 = boundMod;
    } else {
        type = RF_NONE;   // unknown types are dropped
    }
    return true;
}

// loadGeometry (Scene.cpp:191-282)
bool load_geometry(const Node *geometry, SceneBuild &S)
{
    return for_each_element(geometry, [&](const Node *e) {
        rf_primitive o;
        std::memset(&o, 0, sizeof o);
        float boundMod = 1.f;
        if (const char *b = e->attr("isBounding")) boundMod = (std::strcmp(b, "true") == 0) ? -1.f : 1.f;
        int32_t type = RF_NONE;
        if (!classify(e, o.util1, o.util2, type, false, boundMod, S.err)) return false;
        o.type = type;
        if (type != RF_NONE) {
            init_primitive(e, o);
            S.prims.push_back(o);
        }
        return true;
    });
}

// initBV (Scene.cpp:392-452)
bool init_bv(const Node *e, rf_bounding_volume &o, SceneBuild &S)
{
    if (!e->first_child()) return true;
    Mat4 inv;
    bool ok = for_each_element(e, [&](const Node *c) {
        const std::string &t = c->name;
        if (t == "translation") {
            translate(inv, float_attr(c, "x"), float_attr(c, "y"), float_attr(c, "z"), true);
        } else if (t == "scale") {
            scale(inv, float_attr(c, "x"), float_attr(c, "y"), float_attr(c, "z"), true);
        } else if (t == "rotate") {
            rotate(inv, float_attr(c, "degrees"), float_attr(c, "x"), float_attr(c, "y"), float_attr(c, "z"), true);
        } else if (t == "children") {
            o.objStartIndex = (int32_t)S.prims.size();
            if (!load_geometry(c, S)) return false;
            o.objEndIndex = (int32_t)S.prims.size();
        }
        return true;
    });
    if (!ok) return false;
    for (int i = 0; i < 12; ++i) o.invm[i] = inv.m[i];
    return true;
}

// loadBoundingVolumes (Scene.cpp:97-175)
bool load_bvs(const Node *geometry, SceneBuild &S)
{
    return for_each_element(geometry, [&](const Node *e) {
        rf_bounding_volume o;
        std::memset(&o, 0, sizeof o);
        if (e->attr("intermediateCheck")) o.settings = 1;
        int32_t type = RF_NONE;
        if (!classify(e, o.util1, o.util2, type, true, 1.f, S.err)) return false;
        o.type = type;
        if (type != RF_NONE) {
            if (!init_bv(e, o, S)) return false;
            S.bvs.push_back(o);
        }
        return true;
    });
}

int32_t build_scene(const std::string &text, rf_host_scene *out)
{
    std::memset(out, 0, sizeof *out);
    Node doc;
    Parser parser(text);
    if (!parser.parse(doc)) return RF_ERR_PARSE;
    const Node *project = doc.first_child_element("project");
    if (!project) return RF_ERR_PARSE;
    const Node *settings = project->first_child_element("settings");
    const Node *geometry = project->first_child_element("geometry");
    if (!settings || !geometry) return RF_ERR_PARSE;
    // loadSettings (Scene.cpp:64-84); useBoundingVolumes defaults to false
    float globalRayDensity = 0.0f;
    bool useBVs = false;
    for_each_element(settings, [&](const Node *c) {
        if (c->name == "globalRayDensity") globalRayDensity = float_attr(c, "value");
        else if (c->name == "useBoundingVolumes") useBVs = bool_attr(c, "value");
        return true;
    });
    SceneBuild S;
    const bool ok = useBVs ? load_bvs(geometry, S) : load_geometry(geometry, S);
    if (!ok) return RF_ERR_PARSE;
    out->globalRayDensity = globalRayDensity;
    out->useBoundingVolumes = useBVs ? 1 : 0;
    out->n_prims = (int32_t)S.prims.size();
    out->n_bvs = (int32_t)S.bvs.size();
    if (out->n_prims) {
        out->prims = (rf_primitive *)std::malloc(sizeof(rf_primitive) * out->n_prims);
        if (!out->prims) return RF_ERR_OUT_OF_MEMORY;
        std::memcpy(out->prims, S.prims.data(), sizeof(rf_primitive) * out->n_prims);
    }
    if (out->n_bvs) {
        out->bvs = (rf_bounding_volume *)std::malloc(sizeof(rf_bounding_volume) * out->n_bvs);
        if (!out->bvs) { std::free(out->prims); out->prims = nullptr; return RF_ERR_OUT_OF_MEMORY; }
        std::memcpy(out->bvs, S.bvs.data(), sizeof(rf_bounding_volume) * out->n_bvs);
    }
    return RF_OK;
}

}  // namespace

extern "C" int32_t rf_scene_parse_xml(const char *text, size_t len, rf_host_scene *out)
{
    if (!text || !out) return RF_ERR_INVALID_ARG;
    return build_scene(std::string(text, len), out);
}

extern "C" int32_t rf_scene_load_xml(const char *path, rf_host_scene *out)
{
    if (!path || !out) return RF_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof *out);
    std::ifstream f(path, std::ios::binary);
    if (!f) return RF_ERR_IO;   // Scene.cpp:28-34 prints and exits
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return build_scene(text, out);
}

extern "C" void rf_host_scene_free(rf_host_scene *s)
{
    if (!s) return;
    std::free(s->prims);
    std::free(s->bvs);
    std::memset(s, 0, sizeof *s);
}

// output.txt writer (main.cpp:216-251): float accumulation of the int
// hits, division by the double sample count, precision 8, trailing commas.
extern "C" int32_t rf_write_output(const char *path, int32_t n, const int64_t *hits, const double *samples)
{
    if (!path || n < 0 || (n > 0 && (!hits || !samples))) return RF_ERR_INVALID_ARG;
    std::ofstream out(path);
    if (!out) return RF_ERR_IO;
    out.precision(8);
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t j = 0; j < n; ++j) {
            float v = 0.0f;
            v += (float)hits[(size_t)i * n + j];
            v = (float)(v / samples[i]);
            out << v << ",";
        }
        out << std::endl;
    }
    return out ? RF_OK : RF_ERR_IO;
}

extern "C" int32_t rf_normalise(int32_t n, const int64_t *hits, const int64_t *rays, double *vf)
{
    if (n < 0 || (n > 0 && (!hits || !rays || !vf))) return RF_ERR_INVALID_ARG;
    for (int32_t i = 0; i < n; ++i)
        for (int32_t j = 0; j < n; ++j)
            vf[(size_t)i * n + j] = rays[i] > 0 ? (double)hits[(size_t)i * n + j] / (double)rays[i] : 0.0;
    return RF_OK;
}
