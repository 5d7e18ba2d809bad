// MeshIO.cpp -- OpenFOAM and Gmsh ASCII readers, VTU writer.
#include "MeshIO.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <set>
#include <sstream>

namespace cf
{
namespace
{

[[noreturn]] void fail(const std::string& what) { throw Error(what); }

std::string slurp(const std::string& path, const char* who)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) fail(std::string(who) + ": cannot open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

std::string lowercase(std::string s)
{
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

bool allDigits(const std::string& s)
{
    std::size_t i = (!s.empty() && (s[0] == '+' || s[0] == '-')) ? 1 : 0;
    if (i >= s.size()) return false;
    for (; i < s.size(); ++i)
        if (!std::isdigit(static_cast<unsigned char>(s[i]))) return false;
    return true;
}

long strictLong(const std::string& s, const std::string& where, const char* who)
{
    std::size_t used = 0;
    long v = 0;
    try
    {
        v = std::stol(s, &used);
    }
    catch (const std::exception&)
    {
        used = 0;
    }
    if (used != s.size() || s.empty()) fail(std::string(who) + ": bad integer '" + s + "' in " + where);
    return v;
}

double strictDouble(const std::string& s, const std::string& where, const char* who)
{
    std::size_t used = 0;
    double v = 0;
    try
    {
        v = std::stod(s, &used);
    }
    catch (const std::exception&)
    {
        used = 0;
    }
    if (used != s.size() || s.empty()) fail(std::string(who) + ": bad number '" + s + "' in " + where);
    return v;
}

std::string stripQuotes(const std::string& s)
{
    if (s.size() >= 2 && (s.front() == '"' || s.front() == '\'') && s.back() == s.front())
        return s.substr(1, s.size() - 2);
    return s;
}

// ------------------------------------------------------------------ OpenFOAM
// Token stream of the ASCII list syntax: words and the five punctuation
// characters; `//` and `/* */` comments and whitespace are dropped.
enum class Tk : char { Word, LParen, RParen, LBrace, RBrace, Semi };
struct Tok
{
    Tk kind;
    std::string text;
};

std::vector<Tok> scanFoam(const std::string& s, const std::string& path)
{
    std::vector<Tok> out;
    const std::size_t n = s.size();
    std::size_t i = 0;
    auto punct = [](char c) {
        switch (c)
        {
        case '(': return Tk::LParen;
        case ')': return Tk::RParen;
        case '{': return Tk::LBrace;
        case '}': return Tk::RBrace;
        case ';': return Tk::Semi;
        default: return Tk::Word;
        }
    };
    auto comment = [&](std::size_t k) { return k + 1 < n && s[k] == '/' && (s[k + 1] == '/' || s[k + 1] == '*'); };
    while (i < n)
    {
        if (comment(i))
        {
            if (s[i + 1] == '/')
            {
                while (i < n && s[i] != '\n') ++i;
            }
            else
            {
                const std::size_t e = s.find("*/", i + 2);
                if (e == std::string::npos) fail("openfoam: unterminated /* comment in " + path);
                i = e + 2;
            }
            continue;
        }
        const char c = s[i];
        if (std::isspace(static_cast<unsigned char>(c)))
        {
            ++i;
            continue;
        }
        if (punct(c) != Tk::Word)
        {
            out.push_back({punct(c), std::string(1, c)});
            ++i;
            continue;
        }
        std::size_t j = i;
        while (j < n && !std::isspace(static_cast<unsigned char>(s[j])) && punct(s[j]) == Tk::Word && !comment(j))
            ++j;
        out.push_back({Tk::Word, s.substr(i, j - i)});
        i = j;
    }
    return out;
}

// The file's one top-level `N ( ... )` list (brace blocks such as the
// FoamFile header are skipped): count and the token range inside.
struct FoamList
{
    std::vector<Tok> t;
    long count = 0;
    std::size_t pos = 0, end = 0;
    std::string path;

    const Tok& next(const char* what)
    {
        if (pos >= end) fail("openfoam: truncated list (expected " + std::string(what) + ") in " + path);
        return t[pos++];
    }
    void want(Tk k, const char* what)
    {
        if (next(what).kind != k) fail("openfoam: expected " + std::string(what) + " in " + path);
    }
    long integer(const char* what)
    {
        const Tok& k = next(what);
        if (k.kind != Tk::Word || !allDigits(k.text)) fail("openfoam: expected " + std::string(what) + " in " + path);
        return strictLong(k.text, path, "openfoam");
    }
    double number(const char* what)
    {
        const Tok& k = next(what);
        if (k.kind != Tk::Word) fail("openfoam: expected " + std::string(what) + " in " + path);
        return strictDouble(k.text, path, "openfoam");
    }
    void done(const char* what) const
    {
        if (pos != end) fail("openfoam: extra entries in " + std::string(what) + " of " + path);
    }
};

std::size_t closing(const std::vector<Tok>& t, std::size_t open, Tk o, Tk c, const std::string& path)
{
    int depth = 0;
    for (std::size_t i = open; i < t.size(); ++i)
    {
        if (t[i].kind == o) ++depth;
        if (t[i].kind == c && --depth == 0) return i;
    }
    fail("openfoam: unbalanced delimiters in " + path);
}

FoamList foamList(const std::string& path)
{
    const std::string text = slurp(path, "openfoam");
    // a binary mesh says so in its header: format binary;
    const std::string low = lowercase(text);
    const std::size_t f = low.find("format");
    if (f != std::string::npos)
    {
        const std::size_t semi = low.find(';', f);
        if (semi != std::string::npos && low.substr(f, semi - f).find("binary") != std::string::npos)
            fail("openfoam: " + path + " is a binary mesh; write it with ASCII format");
    }
    FoamList L;
    L.path = path;
    L.t = scanFoam(text, path);
    int depth = 0;
    for (std::size_t i = 0; i + 1 < L.t.size(); ++i)
    {
        const Tok& k = L.t[i];
        if (k.kind == Tk::LBrace) ++depth;
        else if (k.kind == Tk::RBrace) --depth;
        else if (depth == 0 && k.kind == Tk::Word && allDigits(k.text) && L.t[i + 1].kind == Tk::LParen)
        {
            L.count = strictLong(k.text, path, "openfoam");
            if (L.count < 0) fail("openfoam: negative list size in " + path);
            L.pos = i + 2;
            L.end = closing(L.t, i + 1, Tk::LParen, Tk::RParen, path);
            return L;
        }
    }
    fail("openfoam: no 'N ( ... )' list in " + path);
}

std::string foamPatchType(const std::string& raw)
{
    const std::string k = lowercase(stripQuotes(raw));
    if (k == "wall" || k == "patch" || k == "67") return "wall";
    if (k == "empty" || k == "symmetryplane" || k == "symmetry") return "symmetry";
    return stripQuotes(raw);
}

// ---------------------------------------------------------------------- Gmsh
std::vector<std::string> words(const std::string& line)
{
    std::vector<std::string> w;
    std::istringstream ss(line);
    for (std::string x; ss >> x;) w.push_back(x);
    return w;
}

long gInt(const std::string& s, const std::string& where)
{
    try
    {
        return std::stol(s);
    }
    catch (const std::exception&)
    {
        fail("gmsh: expected an integer but found '" + s + "' in " + where);
    }
}

double gReal(const std::string& s, const std::string& where)
{
    try
    {
        return std::stod(s);
    }
    catch (const std::exception&)
    {
        fail("gmsh: expected a number but found '" + s + "' in " + where);
    }
}

// non-empty rows of one $Name ... $EndName section, read front to back
struct GSection
{
    std::vector<std::vector<std::string>> rows;
    std::size_t at = 0;
    std::string where;
    bool present() const { return !rows.empty(); }
    bool more() const { return at < rows.size(); }
    const std::vector<std::string>& row()
    {
        if (at >= rows.size())
            fail("gmsh: truncated file: " + where + " ends after " + std::to_string(rows.size()) + " rows");
        return rows[at++];
    }
    // the next row if it is one integer (a node tag line of the 4.x layouts)
    bool tagRow(long& v)
    {
        if (at >= rows.size() || rows[at].size() != 1 || rows[at][0].find_first_of(".eE") != std::string::npos)
            return false;
        v = std::stol(rows[at][0]);
        ++at;
        return true;
    }
};

GSection gSection(const std::vector<std::string>& lines, const std::string& name, const std::string& path,
                  bool required)
{
    const std::string open = "$" + name, close = "$End" + name;
    for (std::size_t i = 0; i < lines.size(); ++i)
    {
        const auto w = words(lines[i]);
        if (w.empty() || w[0] != open) continue;
        GSection s;
        s.where = open + " in " + path;
        for (std::size_t j = i + 1; j < lines.size(); ++j)
        {
            auto r = words(lines[j]);
            if (!r.empty() && r[0] == close) return s;
            if (!r.empty()) s.rows.push_back(std::move(r));
        }
        fail("gmsh: section " + open + " in " + path + " is not terminated");
    }
    if (required) fail("gmsh: " + path + " has no " + open + " section");
    return {};
}

// corner count of first-order Gmsh element types; 0: unsupported here
int cornerCount(int type)
{
    switch (type)
    {
    case 1: return 2;  // line
    case 2: return 3;  // triangle
    case 3: return 4;  // quadrangle
    case 4: return 4;  // tetrahedron
    case 5: return 8;  // hexahedron
    case 6: return 6;  // prism
    case 7: return 5;  // pyramid
    default: return 0;
    }
}

// faces of a volume element as rings of local node numbers (Gmsh numbering:
// hexahedron bottom 0-3, top 4-7; prism bottom 0-2, top 3-5)
const std::vector<std::vector<int>>& faceRings(int type)
{
    static const std::vector<std::vector<int>> tet{{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    static const std::vector<std::vector<int>> hex{{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                                   {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    static const std::vector<std::vector<int>> prism{{0, 1, 2}, {3, 4, 5}, {0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}};
    static const std::vector<std::vector<int>> none;
    return type == 4 ? tet : type == 5 ? hex : type == 6 ? prism : none;
}

// a face by its vertex set (sorted, padded to 5 so triangles and quads differ)
using VertexSet = std::array<label, 5>;
VertexSet vertexSet(const std::vector<label>& v)
{
    VertexSet k;
    k.fill(-1);
    std::vector<label> s(v);
    std::sort(s.begin(), s.end());
    for (std::size_t i = 0; i < s.size() && i < 5; ++i) k[i] = s[i];
    return k;
}

struct GElement
{
    int type = 0;
    long physical = 0;
    std::vector<label> nodes;
};

} // namespace

// ============================================================== OpenFOAM
polyMesh readOpenFoam(const std::string& dirIn, const PatchTypes& patchTypes)
{
    namespace fs = std::filesystem;
    fs::path dir(dirIn);
    if (!fs::exists(dir / "points") && fs::exists(dir / "constant" / "polyMesh" / "points"))
        dir = dir / "constant" / "polyMesh";
    if (!fs::exists(dir / "points"))
        fail("openfoam: no 'points' file in '" + dirIn + "' (nor in its constant/polyMesh subdirectory)");
    const std::string pPath = (dir / "points").string(), fPath = (dir / "faces").string(),
                      oPath = (dir / "owner").string(), nPath = (dir / "neighbour").string(),
                      bPath = (dir / "boundary").string();
    FoamList P = foamList(pPath), F = foamList(fPath), O = foamList(oPath), Nb = foamList(nPath),
             Bd = foamList(bPath);

    std::vector<double> points;
    points.reserve(3 * static_cast<std::size_t>(P.count));
    for (long i = 0; i < P.count; ++i)
    {
        P.want(Tk::LParen, "point '('");
        for (int d = 0; d < 3; ++d) points.push_back(P.number("point coordinate"));
        P.want(Tk::RParen, "point ')'");
    }
    P.done("points");

    std::vector<label> faceStart{0}, faceVerts;
    for (long i = 0; i < F.count; ++i)
    {
        const long nv = F.integer("face vertex count");
        if (nv < 3) fail("openfoam: face " + std::to_string(i) + " has fewer than 3 vertices");
        F.want(Tk::LParen, "face '('");
        for (long k = 0; k < nv; ++k)
        {
            const long v = F.integer("face vertex index");
            if (v < 0) fail("openfoam: negative point index in face " + std::to_string(i));
            faceVerts.push_back(v);
        }
        F.want(Tk::RParen, "face ')'");
        faceStart.push_back(static_cast<label>(faceVerts.size()));
    }
    F.done("faces");

    auto labels = [](FoamList& L, const char* what) {
        std::vector<label> v;
        v.reserve(static_cast<std::size_t>(L.count));
        for (long i = 0; i < L.count; ++i)
        {
            const long x = L.integer(what);
            if (x < 0) fail("openfoam: negative " + std::string(what) + " in " + L.path);
            v.push_back(x);
        }
        L.done(what);
        return v;
    };
    std::vector<label> owner = labels(O, "owner"), neighbour = labels(Nb, "neighbour");

    struct Spec
    {
        std::string name, type;
        long nFaces = 0, start = 0;
    };
    std::vector<Spec> specs;
    for (long i = 0; i < Bd.count; ++i)
    {
        const Tok name = Bd.next("patch name");
        if (name.kind != Tk::Word) fail("openfoam: expected patch name in " + bPath);
        const std::size_t open = Bd.pos;
        Bd.want(Tk::LBrace, "patch '{'");
        const std::size_t close = closing(Bd.t, open, Tk::LBrace, Tk::RBrace, bPath);
        Spec s;
        s.name = stripQuotes(name.text);
        bool type = false, nf = false, sf = false;
        // only type, nFaces and startFace are read; other entries (inGroups ...) are skipped
        for (std::size_t j = Bd.pos; j + 1 < close; ++j)
        {
            if (Bd.t[j].kind != Tk::Word || Bd.t[j + 1].kind != Tk::Word) continue;
            const std::string key = lowercase(Bd.t[j].text);
            if (key == "type")
            {
                s.type = stripQuotes(Bd.t[j + 1].text);
                type = true;
            }
            else if (key == "nfaces")
            {
                s.nFaces = strictLong(Bd.t[j + 1].text, bPath, "openfoam");
                nf = true;
            }
            else if (key == "startface")
            {
                s.start = strictLong(Bd.t[j + 1].text, bPath, "openfoam");
                sf = true;
            }
        }
        if (!type || !nf || !sf) fail("openfoam: patch '" + s.name + "' needs type, nFaces and startFace");
        Bd.pos = close + 1;
        specs.push_back(std::move(s));
    }
    Bd.done("boundary");

    // consistency of the five files
    const label nPoints = static_cast<label>(points.size() / 3);
    const label nFaces = static_cast<label>(faceStart.size()) - 1;
    if (static_cast<label>(owner.size()) != nFaces)
        fail("openfoam: owner has " + std::to_string(owner.size()) + " entries but faces has " +
             std::to_string(nFaces));
    for (label f = 0; f < nFaces; ++f)
        for (label k = faceStart[f]; k < faceStart[f + 1]; ++k)
            if (faceVerts[k] >= nPoints)
                fail("openfoam: face " + std::to_string(f) + " refers to point " + std::to_string(faceVerts[k]) +
                     ", outside 0.." + std::to_string(nPoints - 1));
    label nCells = 0;
    for (label o : owner) nCells = std::max(nCells, o + 1);
    for (label o : neighbour) nCells = std::max(nCells, o + 1);
    if (nCells <= 0) fail("openfoam: mesh " + pPath + " has no cells");
    std::vector<char> owns(static_cast<std::size_t>(nCells), 0);
    for (label o : owner) owns[o] = 1;
    for (label c = 0; c < nCells; ++c)
        if (!owns[c]) fail("openfoam: cell " + std::to_string(c) + " owns no face (owner index out of range?)");
    const label nI = static_cast<label>(neighbour.size());
    if (nI > nFaces) fail("openfoam: more neighbours than faces in " + nPath);
    for (label f = 0; f < nI; ++f)
        if (owner[f] == neighbour[f]) fail("openfoam: face " + std::to_string(f) + " has owner == neighbour");
    label nB = 0;
    for (const Spec& s : specs)
    {
        if (s.nFaces < 0 || s.start < 0 || s.start + s.nFaces > nFaces)
            fail("openfoam: patch '" + s.name + "' range is outside the face list");
        nB += s.nFaces;
    }
    if (nI + nB != nFaces)
        fail("openfoam: neighbour (" + std::to_string(nI) + ") + boundary (" + std::to_string(nB) + ") != faces (" +
             std::to_string(nFaces) + ")");
    std::vector<char> seen(static_cast<std::size_t>(nFaces), 0);
    std::vector<polyMesh::Patch> patches;
    for (const Spec& s : specs)
    {
        if (s.start < nI) fail("openfoam: patch '" + s.name + "' covers internal face " + std::to_string(s.start));
        for (label f = s.start; f < s.start + s.nFaces; ++f)
        {
            if (seen[f]) fail("openfoam: boundary face " + std::to_string(f) + " is in two patches");
            seen[f] = 1;
        }
        const auto it = patchTypes.find(s.name);
        patches.push_back({s.name, it != patchTypes.end() ? it->second : foamPatchType(s.type), s.start, s.nFaces});
    }
    for (label f = nI; f < nFaces; ++f)
        if (!seen[f]) fail("openfoam: boundary face " + std::to_string(f) + " is in no patch");
    return polyMesh(std::move(points), std::move(faceStart), std::move(faceVerts), std::move(owner),
                    std::move(neighbour), std::move(patches));
}

// ================================================================== Gmsh
polyMesh readGmsh(const std::string& path, const PatchTypes& patchTypes)
{
    std::vector<std::string> lines;
    {
        std::istringstream in(slurp(path, "gmsh"));
        for (std::string l; std::getline(in, l);)
        {
            if (!l.empty() && l.back() == '\r') l.pop_back();
            lines.push_back(std::move(l));
        }
    }
    GSection fmt = gSection(lines, "MeshFormat", path, true);
    const auto& fr = fmt.row();
    if (fr.size() < 2) fail("gmsh: malformed $MeshFormat in " + path);
    const double version = gReal(fr[0], "$MeshFormat");
    if (gInt(fr[1], "$MeshFormat") != 0) fail("gmsh: " + path + " is a binary .msh; write it with Mesh.Binary = 0");
    if (!(std::abs(version - 2.2) < 1e-9 || version >= 4.0))
        fail("gmsh: format '" + fr[0] + "' is not supported (ASCII 2.2 and 4.1 are)");
    const bool v4 = version >= 3.0;

    // physical names of surfaces (dimension 2)
    std::map<long, std::string> surfaceName;
    {
        GSection s = gSection(lines, "PhysicalNames", path, false);
        if (s.present())
        {
            const long n = gInt(s.row()[0], "$PhysicalNames");
            for (long i = 0; i < n; ++i)
            {
                const auto& t = s.row();
                if (t.size() < 3) fail("gmsh: malformed $PhysicalNames entry");
                const long dim = gInt(t[0], "$PhysicalNames"), tag = gInt(t[1], "$PhysicalNames");
                std::string name;
                for (std::size_t k = 2; k < t.size(); ++k) name += t[k];
                name.erase(std::remove(name.begin(), name.end(), '"'), name.end());
                if (name.empty()) fail("gmsh: physical group " + std::to_string(tag) + " has no name");
                if (dim == 2) surfaceName[tag] = name;
            }
        }
    }

    // nodes, ordered by tag (untagged nodes take the lowest free tags in file order)
    std::vector<std::pair<long, std::array<double, 3>>> rec;
    long declared = 0;
    {
        GSection s = gSection(lines, "Nodes", path, true);
        if (!s.present()) fail("gmsh: " + path + " has an empty $Nodes section");
        auto xyz = [&](const std::vector<std::string>& t, std::size_t o) {
            if (t.size() < o + 3) fail("gmsh: malformed node coordinate in $Nodes");
            return std::array<double, 3>{gReal(t[o], "$Nodes"), gReal(t[o + 1], "$Nodes"), gReal(t[o + 2], "$Nodes")};
        };
        if (!v4)
        {
            declared = gInt(s.row()[0], "$Nodes");
            for (long i = 0; i < declared; ++i)
            {
                const auto& t = s.row();
                if (t.size() < 4) fail("gmsh: malformed node line in " + path);
                rec.emplace_back(gInt(t[0], "$Nodes"), xyz(t, 1));
            }
        }
        else
        {
            const auto& h = s.row();
            if (h.size() < 4) fail("gmsh: malformed $Nodes header in " + path);
            const long blocks = gInt(h[0], "$Nodes header");
            declared = gInt(h[1], "$Nodes header");
            const long firstTag = gInt(h[2], "$Nodes header");
            for (long b = 0; b < blocks; ++b)
            {
                const auto bh = s.row();
                if (bh.size() < 4) fail("gmsh: malformed $Nodes block header in " + path);
                const long count = gInt(bh[3], "$Nodes block");
                if (count <= 0) continue;
                std::vector<long> tags;
                std::vector<std::array<double, 3>> pts;
                if (bh.size() >= 6)
                {
                    // tag range in the block header; explicit tags after the coordinates when it is not contiguous
                    const long lo = gInt(bh[4], "$Nodes block"), hi = gInt(bh[5], "$Nodes block");
                    for (long i = 0; i < count; ++i) pts.push_back(xyz(s.row(), 0));
                    if (lo != hi)
                        for (long v; static_cast<long>(tags.size()) < count && s.tagRow(v);) tags.push_back(v);
                    if (tags.empty())
                        for (long i = 0; i < count; ++i) tags.push_back(lo + i);
                    (void)firstTag;
                }
                else
                {
                    // Gmsh 4.10+ writes every tag on a line of its own before the coordinates
                    for (long v; static_cast<long>(tags.size()) < count && s.tagRow(v);) tags.push_back(v);
                    const bool tagged = static_cast<long>(tags.size()) == count;
                    for (long i = 0; i < count; ++i) pts.push_back(xyz(s.row(), 0));
                    if (!tagged) tags.assign(static_cast<std::size_t>(count), 0);
                }
                for (std::size_t i = 0; i < pts.size(); ++i) rec.emplace_back(tags[i], pts[i]);
            }
        }
    }
    if (static_cast<long>(rec.size()) != declared)
        fail("gmsh: $Nodes declares " + std::to_string(declared) + " nodes, " + std::to_string(rec.size()) +
             " were read from " + path + " (unsupported node layout?)");
    {
        std::set<long> used;
        for (const auto& r : rec)
            if (r.first > 0) used.insert(r.first);
        long next = 1;
        if (used.size() < rec.size())
            for (auto& r : rec)
            {
                if (r.first > 0) continue;
                while (used.count(next)) ++next;
                r.first = next;
                used.insert(next);
            }
    }
    std::sort(rec.begin(), rec.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::map<long, label> index;
    std::vector<double> points;
    points.reserve(3 * rec.size());
    for (const auto& r : rec)
    {
        if (!index.emplace(r.first, static_cast<label>(points.size() / 3)).second)
            fail("gmsh: duplicate node tag " + std::to_string(r.first) + " in " + path);
        points.insert(points.end(), r.second.begin(), r.second.end());
    }

    // entity -> physical group (format 4): $EntityNames, then the surfaces of $Entities
    std::map<long, long> entityPhysical;
    if (v4)
    {
        GSection en = gSection(lines, "EntityNames", path, false);
        if (en.present())
        {
            const long n = gInt(en.row()[0], "$EntityNames");
            for (long i = 0; i < n; ++i)
            {
                const auto& t = en.row();
                if (t.size() < 4) fail("gmsh: malformed $EntityNames entry");
                if (gInt(t[0], "$EntityNames") == 2)
                    entityPhysical[gInt(t[1], "$EntityNames")] = gInt(t[2], "$EntityNames");
            }
        }
        GSection es = gSection(lines, "Entities", path, false);
        if (es.present())
        {
            const auto h = es.row();
            if (h.size() < 4) fail("gmsh: malformed $Entities header");
            for (int dim = 0; dim < 4; ++dim)
            {
                const long n = gInt(h[dim], "$Entities");
                for (long i = 0; i < n; ++i)
                {
                    const auto& t = es.row();
                    if (dim != 2) continue;
                    if (t.size() < 8) fail("gmsh: malformed $Entities surface line");
                    if (gInt(t[7], "$Entities") > 0)
                    {
                        if (t.size() < 9) fail("gmsh: malformed $Entities surface line");
                        entityPhysical[gInt(t[0], "$Entities")] = gInt(t[8], "$Entities");
                    }
                }
            }
        }
    }

    // elements: volumes (tet, hex, prism) become cells, surfaces name boundary faces
    std::vector<GElement> volumes, surfaces;
    {
        GSection s = gSection(lines, "Elements", path, true);
        if (!s.present()) fail("gmsh: " + path + " has an empty $Elements section");
        auto add = [&](int type, long physical, const std::vector<std::string>& tags) {
            GElement e;
            e.type = type;
            e.physical = physical;
            for (const auto& tok : tags)
            {
                const long tag = gInt(tok, "$Elements");
                const auto it = index.find(tag);
                if (it == index.end())
                    fail("gmsh: an element refers to node " + std::to_string(tag) + ", which $Nodes does not define");
                e.nodes.push_back(it->second);
            }
            for (std::size_t k = 1; k < e.nodes.size(); ++k)
                if (e.nodes[k] == e.nodes[k - 1]) fail("gmsh: element " + std::to_string(type) + " repeats a node");
            if (type == 2 || type == 3) surfaces.push_back(std::move(e));
            else if (type >= 4 && type <= 6) volumes.push_back(std::move(e));
            else if (type == 7) fail("gmsh: pyramid elements are not supported");
        };
        auto checkType = [&](int type) {
            if (cornerCount(type) <= 0)
                fail("gmsh: element type " + std::to_string(type) + " in " + path +
                     " is not a first-order element; mesh with Linear order");
            return cornerCount(type);
        };
        if (!v4)
        {
            const long n = gInt(s.row()[0], "$Elements");
            for (long i = 0; i < n; ++i)
            {
                const auto& t = s.row();
                if (t.size() < 3) fail("gmsh: malformed element line in " + path);
                const int type = static_cast<int>(gInt(t[1], "$Elements"));
                const int nc = checkType(type);
                const long nTags = gInt(t[2], "$Elements");
                if (static_cast<long>(t.size()) < 3 + nTags + nc) fail("gmsh: truncated element line in " + path);
                std::vector<std::string> tags(t.begin() + 3 + nTags, t.begin() + 3 + nTags + nc);
                add(type, nTags > 0 ? gInt(t[3], "$Elements") : 0, tags);
            }
        }
        else
        {
            const auto h = s.row();
            if (h.size() < 4) fail("gmsh: malformed $Elements header in " + path);
            const long blocks = gInt(h[0], "$Elements header"), declaredEl = gInt(h[1], "$Elements header");
            long read = 0;
            for (long b = 0; b < blocks; ++b)
            {
                const auto bh = s.row();
                if (bh.size() < 4) fail("gmsh: malformed $Elements block header in " + path);
                const long entity = gInt(bh[1], "$Elements block");
                const int type = static_cast<int>(gInt(bh[2], "$Elements block"));
                const long count = gInt(bh[3], "$Elements block");
                const int nc = checkType(type);
                const auto pit = entityPhysical.find(entity);
                const long physical = pit != entityPhysical.end() ? pit->second : 0;
                bool withTag = false;
                for (long i = 0; i < count; ++i, ++read)
                {
                    const auto& t = s.row();
                    const long w = static_cast<long>(t.size());
                    if (i == 0)
                    {
                        // the element tag column is optional: the block's first line decides
                        if (w == nc + 1) withTag = true;
                        else if (w != nc)
                            fail("gmsh: element line with " + std::to_string(w) + " entries, expected " +
                                 std::to_string(nc) + " or " + std::to_string(nc + 1));
                    }
                    else if (w != nc + (withTag ? 1 : 0))
                        fail("gmsh: inconsistent element line width in " + path);
                    add(type, physical, std::vector<std::string>(t.begin() + (withTag ? 1 : 0), t.end()));
                }
            }
            if (read != declaredEl)
                fail("gmsh: $Elements declares " + std::to_string(declaredEl) + " elements, " +
                     std::to_string(read) + " were read from " + path + " (unsupported element layout?)");
        }
    }
    if (volumes.empty()) fail("gmsh: " + path + " has no volume elements");

    // faces from the cells, first appearance wins the vertex order
    struct Face
    {
        std::vector<label> v;
        label a = -1, b = -1;
    };
    std::vector<Face> faces;
    std::map<VertexSet, std::size_t> where;
    for (label c = 0; c < static_cast<label>(volumes.size()); ++c)
        for (const auto& ring : faceRings(volumes[c].type))
        {
            std::vector<label> v;
            for (int k : ring) v.push_back(volumes[c].nodes[k]);
            const auto ins = where.emplace(vertexSet(v), faces.size());
            if (ins.second)
            {
                faces.push_back({std::move(v), c, -1});
                continue;
            }
            Face& f = faces[ins.first->second];
            if (f.b != -1)
                fail("gmsh: a face belongs to three cells (cell " + std::to_string(f.a) + "/" + std::to_string(f.b) +
                     "/" + std::to_string(c) + "): the mesh is not manifold");
            f.b = c;
        }

    // boundary face names: the named surface element with the same vertex set
    // (a quad face also matches any triangle of its four corners)
    std::map<VertexSet, std::string> named;
    for (const auto& e : surfaces)
    {
        const auto it = surfaceName.find(e.physical);
        if (it != surfaceName.end()) named.emplace(vertexSet(e.nodes), it->second);
    }
    auto surfaceOf = [&](const std::vector<label>& v) -> std::string {
        auto it = named.find(vertexSet(v));
        if (it != named.end()) return it->second;
        if (v.size() == 4)
            for (int skip = 0; skip < 4; ++skip)
            {
                std::vector<label> tri;
                for (int k = 0; k < 4; ++k)
                    if (k != skip) tri.push_back(v[k]);
                it = named.find(vertexSet(tri));
                if (it != named.end()) return it->second;
            }
        return {};
    };
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = points[d];
    for (std::size_t i = 0; i < points.size() / 3; ++i)
        for (int d = 0; d < 3; ++d)
        {
            lo[d] = std::min(lo[d], points[3 * i + d]);
            hi[d] = std::max(hi[d], points[3 * i + d]);
        }
    const double eps = 1e-9 * std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
    static const char* plane[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
    auto planeOf = [&](const std::vector<label>& v) -> std::string {
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side)
            {
                const double x = side ? hi[d] : lo[d];
                bool on = true;
                for (label p : v)
                    if (std::abs(points[3 * p + d] - x) > eps)
                    {
                        on = false;
                        break;
                    }
                if (on) return plane[2 * d + side];
            }
        return {};
    };

    std::vector<label> faceStart{0}, faceVerts, owner, neighbour;
    for (Face& f : faces)
    {
        if (f.b < 0) continue;
        if (f.a > f.b) std::swap(f.a, f.b);
        faceVerts.insert(faceVerts.end(), f.v.begin(), f.v.end());
        faceStart.push_back(static_cast<label>(faceVerts.size()));
        owner.push_back(f.a);
        neighbour.push_back(f.b);
    }
    std::vector<std::string> order;  // patch names in order of first appearance
    std::map<std::string, std::vector<const Face*>> members;
    for (const Face& f : faces)
    {
        if (f.b >= 0) continue;
        std::string name = named.empty() ? std::string() : surfaceOf(f.v);
        if (name.empty()) name = planeOf(f.v);
        if (name.empty())
        {
            std::string vs;
            for (label p : f.v) vs += " " + std::to_string(p);
            fail("gmsh: boundary face with nodes" + vs +
                 " belongs to no physical group and lies on no bounding-box plane; give every boundary surface "
                 "a physical group in Gmsh");
        }
        auto& m = members[name];
        if (m.empty()) order.push_back(name);
        m.push_back(&f);
    }
    std::vector<polyMesh::Patch> patches;
    for (const std::string& name : order)
    {
        const auto it = patchTypes.find(name);
        polyMesh::Patch p{name, it != patchTypes.end() ? it->second : "wall", static_cast<label>(owner.size()),
                          static_cast<label>(members[name].size())};
        for (const Face* f : members[name])
        {
            faceVerts.insert(faceVerts.end(), f->v.begin(), f->v.end());
            faceStart.push_back(static_cast<label>(faceVerts.size()));
            owner.push_back(f->a);
        }
        patches.push_back(std::move(p));
    }
    return polyMesh(std::move(points), std::move(faceStart), std::move(faceVerts), std::move(owner),
                    std::move(neighbour), std::move(patches));
}

// =================================================================== VTU
void writeVtu(const std::string& path, const polyMesh& mesh, const std::vector<VtuField>& fields)
{
    const label nC = mesh.nCells(), nF = mesh.nFaces(), nI = mesh.nInternalFaces();
    for (const VtuField& f : fields)
    {
        if (f.components != 1 && f.components != 3) fail("vtk: field '" + f.name + "' must have 1 or 3 components");
        if (!f.data) fail("vtk: field '" + f.name + "' has no data");
    }
    // cell -> faces in face order (owner side, then neighbour side as the face comes)
    std::vector<label> start(static_cast<std::size_t>(nC) + 1, 0), list;
    for (label f = 0; f < nF; ++f)
    {
        ++start[mesh.owner()[f] + 1];
        if (f < nI) ++start[mesh.neighbour()[f] + 1];
    }
    for (label c = 0; c < nC; ++c) start[c + 1] += start[c];
    list.resize(static_cast<std::size_t>(start[nC]));
    {
        std::vector<label> fill(start.begin(), start.end() - 1);
        for (label f = 0; f < nF; ++f)
        {
            list[fill[mesh.owner()[f]]++] = f;
            if (f < nI) list[fill[mesh.neighbour()[f]]++] = f;
        }
    }
    FILE* out = std::fopen(path.c_str(), "w");
    if (!out) fail("vtk: cannot open '" + path + "' for writing");
    const auto& P = mesh.points();
    const auto& fs = mesh.faceStart();
    const auto& fv = mesh.faceVerts();
    std::fprintf(out,
                 "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
                 "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n<UnstructuredGrid>\n"
                 "<Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n",
                 static_cast<long long>(mesh.nPoints()), static_cast<long long>(nC));
    std::fprintf(out, "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (label p = 0; p < mesh.nPoints(); ++p)
        std::fprintf(out, "%.12g %.12g %.12g\n", P[3 * p], P[3 * p + 1], P[3 * p + 2]);
    std::fprintf(out, "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" "
                      "format=\"ascii\">\n");
    std::vector<long long> offsets, faceOffsets;
    long long off = 0, foff = 0;
    for (label c = 0; c < nC; ++c)
    {
        // the cell's distinct vertices in order of first appearance over its faces
        std::vector<label> verts;
        for (label k = start[c]; k < start[c + 1]; ++k)
            for (label j = fs[list[k]]; j < fs[list[k] + 1]; ++j)
                if (std::find(verts.begin(), verts.end(), fv[j]) == verts.end()) verts.push_back(fv[j]);
        for (label v : verts) std::fprintf(out, "%lld ", static_cast<long long>(v));
        std::fprintf(out, "\n");
        off += static_cast<long long>(verts.size());
        offsets.push_back(off);
    }
    std::fprintf(out, "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
    for (long long o : offsets) std::fprintf(out, "%lld ", o);
    std::fprintf(out, "\n</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
    for (label c = 0; c < nC; ++c) std::fprintf(out, "42 ");
    std::fprintf(out, "\n</DataArray>\n<DataArray type=\"Int64\" Name=\"faces\" format=\"ascii\">\n");
    for (label c = 0; c < nC; ++c)
    {
        // face stream of a polyhedron: nFaces, then per face its vertex count and vertices
        std::fprintf(out, "%lld ", static_cast<long long>(start[c + 1] - start[c]));
        foff += 1;
        for (label k = start[c]; k < start[c + 1]; ++k)
        {
            const label f = list[k];
            std::fprintf(out, "%lld ", static_cast<long long>(fs[f + 1] - fs[f]));
            for (label j = fs[f]; j < fs[f + 1]; ++j) std::fprintf(out, "%lld ", static_cast<long long>(fv[j]));
            foff += 1 + (fs[f + 1] - fs[f]);
        }
        std::fprintf(out, "\n");
        faceOffsets.push_back(foff);
    }
    std::fprintf(out, "</DataArray>\n<DataArray type=\"Int64\" Name=\"faceoffsets\" format=\"ascii\">\n");
    for (long long o : faceOffsets) std::fprintf(out, "%lld ", o);
    std::fprintf(out, "\n</DataArray>\n</Cells>\n");
    if (!fields.empty())
    {
        std::fprintf(out, "<CellData>\n");
        for (const VtuField& f : fields)
        {
            if (f.components == 1)
            {
                std::fprintf(out, "<DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", f.name.c_str());
                for (label c = 0; c < nC; ++c) std::fprintf(out, "%.12g\n", f.data[c]);
            }
            else
            {
                std::fprintf(out, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n",
                             f.name.c_str());
                for (label c = 0; c < nC; ++c)
                    std::fprintf(out, "%.12g %.12g %.12g\n", f.data[3 * c], f.data[3 * c + 1], f.data[3 * c + 2]);
            }
            std::fprintf(out, "</DataArray>\n");
        }
        std::fprintf(out, "</CellData>\n");
    }
    std::fprintf(out, "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
    const bool bad = std::ferror(out) != 0;
    std::fclose(out);
    if (bad) fail("vtk: write error on '" + path + "'");
}

PatchTypes parsePatchTypes(const std::string& spec)
{
    PatchTypes m;
    std::size_t i = 0;
    while (i < spec.size())
    {
        std::size_t j = spec.find(';', i);
        if (j == std::string::npos) j = spec.size();
        const std::string item = spec.substr(i, j - i);
        if (!item.empty())
        {
            const std::size_t eq = item.find('=');
            if (eq == std::string::npos || eq == 0 || eq + 1 == item.size())
                fail("patch types: expected name=type, got '" + item + "'");
            m[item.substr(0, eq)] = item.substr(eq + 1);
        }
        i = j + 1;
    }
    return m;
}

} // namespace cf
