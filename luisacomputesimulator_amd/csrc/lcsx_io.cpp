// Scene JSON and OBJ input (reference Application/app_simulation_demo_config.cpp:49-766,
// Solver/MeshOperation/mesh_reader.cpp:355-541). A small self-contained JSON reader: objects,
// arrays, numbers, strings, booleans and null.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include "lcsx_io.h"

namespace lcsx
{

namespace json
{
struct Value
{
    enum Kind
    {
        Null,
        Bool,
        Num,
        Str,
        Arr,
        Obj
    } kind = Null;
    bool                                         b = false;
    double                                       n = 0.0;
    std::string                                  s;
    std::vector<Value>                           a;
    std::vector<std::pair<std::string, Value>>   o;
    const Value* get(const std::string& k) const
    {
        for (auto& kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct Parser
{
    const std::string& t;
    size_t             i = 0;
    explicit Parser(const std::string& text) : t(text) {}
    void ws()
    {
        while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) i++;
    }
    bool fail(const char* m)
    {
        throw std::runtime_error(std::string("JSON: ") + m + " at offset " + std::to_string(i));
    }
    Value parse()
    {
        ws();
        if (i >= t.size()) fail("unexpected end");
        char  c = t[i];
        Value v;
        if (c == '{')
        {
            v.kind = Value::Obj;
            i++;
            ws();
            if (t[i] == '}')
            {
                i++;
                return v;
            }
            while (true)
            {
                ws();
                Value k = parse();
                if (k.kind != Value::Str) fail("object key must be a string");
                ws();
                if (t[i] != ':') fail("expected ':'");
                i++;
                v.o.emplace_back(k.s, parse());
                ws();
                if (t[i] == ',') { i++; continue; }
                if (t[i] == '}') { i++; break; }
                fail("expected ',' or '}'");
            }
        }
        else if (c == '[')
        {
            v.kind = Value::Arr;
            i++;
            ws();
            if (t[i] == ']')
            {
                i++;
                return v;
            }
            while (true)
            {
                v.a.push_back(parse());
                ws();
                if (t[i] == ',') { i++; continue; }
                if (t[i] == ']') { i++; break; }
                fail("expected ',' or ']'");
            }
        }
        else if (c == '"')
        {
            v.kind = Value::Str;
            i++;
            while (i < t.size() && t[i] != '"')
            {
                if (t[i] == '\\' && i + 1 < t.size()) i++;
                v.s.push_back(t[i++]);
            }
            i++;
        }
        else if (t.compare(i, 4, "true") == 0) { v.kind = Value::Bool; v.b = true; i += 4; }
        else if (t.compare(i, 5, "false") == 0) { v.kind = Value::Bool; v.b = false; i += 5; }
        else if (t.compare(i, 4, "null") == 0) { i += 4; }
        else
        {
            v.kind       = Value::Num;
            const char* s = t.c_str() + i;
            char*       e = nullptr;
            v.n          = std::strtod(s, &e);
            if (e == s) fail("bad value");
            i += (size_t)(e - s);
        }
        return v;
    }
};
}  // namespace json

bool read_obj(const std::string& path, std::vector<float>& verts, std::vector<int32_t>& tris, std::string& err)
{
    std::ifstream f(path);
    if (!f)
    {
        err = "cannot open " + path;
        return false;
    }
    std::string line;
    while (std::getline(f, line))
    {
        std::istringstream ss(line);
        std::string        tag;
        ss >> tag;
        if (tag == "v")
        {
            float a, b, c;
            ss >> a >> b >> c;
            verts.push_back(a);
            verts.push_back(b);
            verts.push_back(c);
        }
        else if (tag == "f")
        {
            std::vector<int32_t> poly;
            std::string          tok;
            while (ss >> tok)
            {
                int idx = std::atoi(tok.c_str());
                if (idx < 0) idx = (int)(verts.size() / 3) + idx + 1;
                poly.push_back(idx - 1);
            }
            for (size_t k = 1; k + 1 < poly.size(); k++)
            {
                tris.push_back(poly[0]);
                tris.push_back(poly[k]);
                tris.push_back(poly[k + 1]);
            }
        }
    }
    return true;
}

namespace
{
const std::map<std::string, int> kMethods = {
    {"None", 0},       {"FromIndices", 1}, {"FromFunction", 2}, {"Left", 3},        {"Right", 4},     {"Front", 5},
    {"Back", 6},       {"Up", 7},          {"Down", 8},         {"LeftUp", 9},      {"LeftDown", 10}, {"LeftFront", 11},
    {"LeftBack", 12},  {"RightUp", 13},    {"RightDown", 14},   {"RightFront", 15}, {"RightBack", 16}, {"FrontUp", 17},
    {"FrontDown", 18}, {"BackUp", 19},     {"BackDown", 20},    {"All", 21}};

bool num(const json::Value* v, float& out)
{
    if (v && v->kind == json::Value::Num)
    {
        out = (float)v->n;
        return true;
    }
    return false;
}
bool vec3(const json::Value* v, float out[3])
{
    if (!v || v->kind != json::Value::Arr || v->a.size() < 3) return false;
    for (int k = 0; k < 3; k++) out[k] = v->a[k].kind == json::Value::Num ? (float)v->a[k].n : 0.0f;
    return true;
}

// Model matrix translate(t) * Rx * Ry * Rz * scale(s) (reference Core/affine_position.h:41-98):
// scale first, then rotate about z, y and x, then translate.
void model_transform(const float tr[3], const float rot[3], const float sc[3], std::vector<float>& x)
{
    float cx = std::cos(rot[0]), sx = std::sin(rot[0]), cy = std::cos(rot[1]), sy = std::sin(rot[1]);
    float cz = std::cos(rot[2]), sz = std::sin(rot[2]);
    for (size_t i = 0; i < x.size() / 3; i++)
    {
        float p[3] = {x[3 * i] * sc[0], x[3 * i + 1] * sc[1], x[3 * i + 2] * sc[2]};
        float x1 = cz * p[0] - sz * p[1], y1 = sz * p[0] + cz * p[1], z1 = p[2];
        float x2 = cy * x1 + sy * z1, y2 = y1, z2 = -sy * x1 + cy * z1;
        float y3 = cx * y2 - sx * z2, z3 = sx * y2 + cx * z2;
        x[3 * i]     = x2 + tr[0];
        x[3 * i + 1] = y3 + tr[1];
        x[3 * i + 2] = z3 + tr[2];
    }
}
}  // namespace

bool load_scene_json(const std::string& path, const std::string& mesh_dir, lcsx_params& prm, std::vector<MeshDesc>& meshes,
                     std::string& err)
{
    std::ifstream f(path);
    if (!f)
    {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream buf;
    buf << f.rdbuf();
    std::string text = buf.str();
    json::Value root;
    try
    {
        json::Parser p(text);
        root = p.parse();
    }
    catch (const std::exception& e)
    {
        err = e.what();
        return false;
    }
    if (root.kind != json::Value::Obj)
    {
        err = "scene root is not an object";
        return false;
    }
    auto get_bool = [&](const char* k, int& out) {
        const json::Value* v = root.get(k);
        if (v && v->kind == json::Value::Bool) out = v->b ? 1 : 0;
    };
    get_bool("use_floor", prm.use_floor);
    get_bool("use_self_collision", prm.use_self_collision);
    get_bool("use_ccd_linesearch", prm.use_ccd_linesearch);
    get_bool("use_energy_linesearch", prm.use_energy_linesearch);
    num(root.get("implicit_dt"), prm.implicit_dt);
    num(root.get("stiffness_collision"), prm.stiffness_collision);
    num(root.get("stiffness_dirichlet"), prm.stiffness_dirichlet);
    float tmp;
    if (num(root.get("nonlinear_iter_count"), tmp)) prm.nonlinear_iter_count = (int)tmp;
    if (num(root.get("pcg_iter_count"), tmp)) prm.pcg_iter_count = (int)tmp;
    if (num(root.get("contact_energy_type"), tmp))
    {
        if (tmp != 0.0f && tmp != 1.0f)
        {
            err = "contact_energy_type must be 0 (quadratic) or 1 (log barrier)";
            return false;
        }
        prm.contact_energy_type = (int)tmp;
    }
    vec3(root.get("gravity"), prm.gravity);
    const json::Value* shells = root.get("shells");
    if (!shells || shells->kind != json::Value::Arr) return true;
    for (const json::Value& sh : shells->a)
    {
        if (sh.kind != json::Value::Obj) continue;
        // shell_type and material.type (app_simulation_demo_config.cpp:362-375, 397-420, 540-575):
        // Cloth (also "stretch_model") or Rigid; the material's "type", when given, must agree
        const json::Value* st    = sh.get("shell_type");
        std::string        stype = st && st->kind == json::Value::Str ? st->s : "Cloth";
        if (stype == "stretch_model") stype = "Cloth";
        const json::Value* mat = sh.get("material");
        if (mat && mat->kind == json::Value::Arr && !mat->a.empty()) mat = &mat->a[0];
        const json::Value* mt = mat && mat->kind == json::Value::Obj ? mat->get("type") : nullptr;
        if (mt && mt->kind == json::Value::Str)
        {
            if (!st) stype = mt->s;
            else if (mt->s != stype)
            {
                err = "shell_type " + stype + " and material type " + mt->s + " disagree";
                return false;
            }
        }
        if (stype != "Cloth" && stype != "Rigid")
        {
            err = "shell_type " + stype + " is not supported by this build (Cloth and Rigid)";
            return false;
        }
        MeshDesc m;
        lcsx_material_default(&m.mat);
        if (stype == "Rigid") m.mat.body = 1;  // RigidMaterial: is_shell defaults to true
        const json::Value* mn = sh.get("model_name");
        if (!mn || mn->kind != json::Value::Str)
        {
            err = "shell without model_name";
            return false;
        }
        std::string file = mn->s;
        std::string cand = mesh_dir.empty() ? file : mesh_dir + "/" + file;
        std::ifstream test(cand);
        if (!test)
        {
            size_t pos = file.find_last_of("/\\");
            cand       = mesh_dir + "/" + (pos == std::string::npos ? file : file.substr(pos + 1));
        }
        if (!read_obj(cand, m.x, m.tris, err)) return false;
        float tr[3] = {0, 0, 0}, rot[3] = {0, 0, 0}, sc[3] = {1, 1, 1};
        vec3(sh.get("translation"), tr);
        vec3(sh.get("rotation"), rot);
        const json::Value* s = sh.get("scale");
        if (s && s->kind == json::Value::Num) sc[0] = sc[1] = sc[2] = (float)s->n;
        else vec3(s, sc);
        std::vector<float> model = m.x;  // pins are selected on model coordinates
        model_transform(tr, rot, sc, m.x);
        if (mat && mat->kind == json::Value::Obj)
        {
            num(mat->get("thickness"), m.mat.thickness);
            num(mat->get("mass"), m.mat.mass);
            num(mat->get("density"), m.mat.density);
            num(mat->get("d_hat"), m.mat.d_hat);
            if (stype == "Rigid")
            {
                num(mat->get("stiffness"), m.mat.rigid_stiffness);
                const json::Value* isv = mat->get("is_shell");
                if (isv && isv->kind == json::Value::Bool && !isv->b) m.mat.body = 2;  // rigid solid
            }
            else
            {
                num(mat->get("youngs_modulus"), m.mat.youngs_modulus);
                num(mat->get("poisson_ratio"), m.mat.poisson_ratio);
                num(mat->get("area_bending_stiffness"), m.mat.area_bending_stiffness);
                // stretch_model: "Spring" or the enum value 0 selects springs, anything else FEM_BW98
                // (app_simulation_demo_config.cpp:455-475)
                const json::Value* sm = mat->get("stretch_model");
                if (sm && sm->kind == json::Value::Str) m.mat.stretch_spring = sm->s == "Spring";
                else if (sm && sm->kind == json::Value::Num) m.mat.stretch_spring = (int)sm->n == 0;
            }
        }
        const json::Value* fps = sh.get("fixed_points");
        if (fps && fps->kind == json::Value::Arr)
        {
            MeshDesc sel;
            sel.x = model;
            for (const json::Value& fp : fps->a)
            {
                int                method = 21;
                const json::Value* mm     = fp.get("method");
                if (mm && mm->kind == json::Value::Str)
                {
                    auto it = kMethods.find(mm->s);
                    method  = it == kMethods.end() ? 21 : it->second;
                }
                float              range = 0.001f;
                const json::Value* rg    = fp.get("range");
                if (rg && rg->kind == json::Value::Arr && !rg->a.empty()) range = (float)rg->a[0].n;
                lcsx_fixed_anim an{};
                an.scale[0] = an.scale[1] = an.scale[2] = 1.0f;
                an.rot_axis[0] = 1.0f;
                const json::Value* fi = fp.get("fixed_info");
                if (fi && fi->kind == json::Value::Obj)
                {
                    auto gb = [&](const char* k, int& out) {
                        const json::Value* v = fi->get(k);
                        if (v && v->kind == json::Value::Bool) out = v->b ? 1 : 0;
                    };
                    gb("use_translate", an.use_translate);
                    gb("use_rotate", an.use_rotate);
                    gb("use_scale", an.use_scale);
                    vec3(fi->get("translate"), an.translate);
                    vec3(fi->get("rotCenter"), an.rot_center);
                    vec3(fi->get("rotAxis"), an.rot_axis);
                    num(fi->get("rotAngVelDeg"), an.rot_deg_per_s);
                    if (vec3(fi->get("scale"), an.scale)) an.use_scale = 1;
                }
                pin_by_method(sel, method, range, an);
            }
            m.pinned = sel.pinned;
            m.anim   = sel.anim;
        }
        meshes.push_back(std::move(m));
    }
    return true;
}

}  // namespace lcsx
