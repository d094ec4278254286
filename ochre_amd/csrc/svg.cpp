// SVG ingestion: a document -> the draw calls the reference's demo makes.
//
// The reference demo (examples/svg.rs:103-164) parses a file with usvg 0.13
// (a crates.io dev-dependency, not vendored: Cargo.toml:14-17) and walks the
// tree: for every path node, in document order, a fill with a solid colour
// becomes Rasterizer::new + fill(path, transform) + finish, then a stroke
// with a solid colour becomes Rasterizer::new + stroke(path, width,
// transform) + finish (svg.rs:140-156).  The path is usvg's PathData (MoveTo,
// LineTo, CurveTo, ClosePath in absolute f64 user units) cast to f32, and the
// transform is the node's own usvg transform (svg.rs:107-111).
//
// usvg 0.13 behaviour restated here (its published conversion rules):
//   * path data (svgtypes grammar): relative commands are made absolute in
//     f64; H/V become LineTo; S reflects the previous C/S second control
//     point, else uses the current point; Q/T become CurveTo by degree
//     elevation ((p0 + 2 c) / 3, (p + 2 c) / 3); A becomes cubics (below) or
//     a LineTo when the radii are degenerate; repeated Z are dropped; after
//     Z the current point is the subpath start and a relative m is taken from
//     it; the first parse error ends the data (the segments before it stay);
//     data not starting with a MoveTo is empty; a path of fewer than two
//     segments is dropped.
//   * groups: a group without opacity, clip-path, mask or filter is removed
//     and its transform prepended to its children, so a path's transform is
//     the product of the transforms from the nearest kept group down to the
//     path (svg.rs uses node.transform(), the transform relative to the
//     parent node; a kept group's own transform is not applied by the demo).
//     The root viewBox is not a transform in usvg's tree and the demo does
//     not apply it.
//   * paint: fill and stroke are inherited presentation attributes (a
//     `style` declaration overrides the attribute); fill defaults to black,
//     stroke to none; "none" disables; a url(#...) paint is not a colour (the
//     demo skips it); stroke-width defaults to 1 and a width <= 0 disables the
//     stroke; fill-opacity / stroke-opacity default to 1.
//   * shapes: rect (with rounded corners as arcs), circle, ellipse, line,
//     polyline, polygon, converted to path data as usvg's shapes module does.
//   * not rendered: defs, symbol, clipPath, mask, pattern, marker, text (usvg
//     needs fonts), image, and any element with display="none".
// Arcs follow kurbo's published Arc::from_svg_arc / append_iter (SVG 1.1
// implementation notes F.6.5 centre parameterisation; cubic segments with
// arms (4/3) tan(step / 4), the step count from the tolerance 0.1).  None of
// the three documents the reference ships uses an arc, a shape element, a
// rotation or a kept group; for those parts parity with usvg is unpinned.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ochre.h"

namespace {

struct Affine {  // SVG matrix(a b c d e f): x' = a x + c y + e, y' = b x + d y + f
    double a = 1, b = 0, c = 0, d = 1, e = 0, f = 0;
};

// m1 * m2 (usvg's Transform::append: self = self * other).
Affine mul(const Affine& m1, const Affine& m2) {
    Affine r;
    r.a = m1.a * m2.a + m1.c * m2.b;
    r.b = m1.b * m2.a + m1.d * m2.b;
    r.c = m1.a * m2.c + m1.c * m2.d;
    r.d = m1.b * m2.c + m1.d * m2.d;
    r.e = m1.a * m2.e + m1.c * m2.f + m1.e;
    r.f = m1.b * m2.e + m1.d * m2.f + m1.f;
    return r;
}

bool is_identity(const Affine& m) {
    return m.a == 1 && m.b == 0 && m.c == 0 && m.d == 1 && m.e == 0 && m.f == 0;
}

// ------------------------------------------------------------ lexing

bool is_ws(char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r'; }

struct Cursor {
    const char* p;
    const char* end;
    bool done() const { return p >= end; }
    char peek() const { return p < end ? *p : '\0'; }
    void skip_ws() {
        while (p < end && is_ws(*p)) p++;
    }
    void skip_ws_comma() {
        skip_ws();
        if (p < end && *p == ',') {
            p++;
            skip_ws();
        }
    }
};

// An SVG number (svgtypes grammar): [sign] (digits [. digits] | . digits)
// [(e|E) [sign] digits].  The exponent is taken only when digits follow, so
// "1em" is the number 1 and the unit "em".  The value is the correctly
// rounded f64 of the text (strtod).
bool parse_number(Cursor& c, double* out) {
    const char* s = c.p;
    const char* q = s;
    if (q < c.end && (*q == '+' || *q == '-')) q++;
    const char* digits = q;
    while (q < c.end && std::isdigit((unsigned char)*q)) q++;
    bool any = q > digits;
    if (q < c.end && *q == '.') {
        const char* frac = ++q;
        while (q < c.end && std::isdigit((unsigned char)*q)) q++;
        any = any || q > frac;
    }
    if (!any) return false;
    if (q < c.end && (*q == 'e' || *q == 'E')) {
        const char* r = q + 1;
        if (r < c.end && (*r == '+' || *r == '-')) r++;
        if (r < c.end && std::isdigit((unsigned char)*r)) {
            while (r < c.end && std::isdigit((unsigned char)*r)) r++;
            q = r;
        }
    }
    std::string tmp(s, q);
    *out = std::strtod(tmp.c_str(), nullptr);
    c.p = q;
    return true;
}

// A list number: the number, then optional white space and one comma.
bool list_number(Cursor& c, double* out) {
    c.skip_ws();
    if (!parse_number(c, out)) return false;
    c.skip_ws_comma();
    return true;
}

// An arc flag: a single '0' or '1' (no separator needed after it).
bool list_flag(Cursor& c, bool* out) {
    c.skip_ws();
    if (c.done() || (c.peek() != '0' && c.peek() != '1')) return false;
    *out = c.peek() == '1';
    c.p++;
    c.skip_ws_comma();
    return true;
}

// ------------------------------------------------------------ path data

enum Seg { S_MOVE, S_LINE, S_CURVE, S_CLOSE };
struct Segment {
    Seg kind;
    double x1, y1, x2, y2, x, y;
};

struct PathData {
    std::vector<Segment> segs;
    void move_to(double x, double y) { segs.push_back(Segment{S_MOVE, 0, 0, 0, 0, x, y}); }
    void line_to(double x, double y) { segs.push_back(Segment{S_LINE, 0, 0, 0, 0, x, y}); }
    void curve_to(double x1, double y1, double x2, double y2, double x, double y) {
        segs.push_back(Segment{S_CURVE, x1, y1, x2, y2, x, y});
    }
    void close() { segs.push_back(Segment{S_CLOSE, 0, 0, 0, 0, 0, 0}); }
    // the current point: the last segment's end point (a Close's is the
    // subpath start, tracked by the caller)
    bool last_pos(double* x, double* y) const {
        for (size_t i = segs.size(); i-- > 0;) {
            if (segs[i].kind != S_CLOSE) {
                *x = segs[i].x;
                *y = segs[i].y;
                return true;
            }
        }
        return false;
    }
    // a quadratic by degree elevation from the current point
    void quad_to(double px, double py, double x1, double y1, double x, double y) {
        curve_to((px + 2.0 * x1) / 3.0, (py + 2.0 * y1) / 3.0, (x + 2.0 * x1) / 3.0, (y + 2.0 * y1) / 3.0, x, y);
    }
    void arc_to(double px, double py, double rx, double ry, double x_rot_deg, bool large, bool sweep, double x,
                double y);
};

// Elliptical arc from (px, py) to (x, y) as cubics: the endpoint-to-centre
// conversion of the SVG implementation notes (F.6.5, F.6.6 radii scaling),
// then kurbo's subdivision: n = ceil(max(1.1163 * max(rx, ry) / tol)^(1/6),
// 3.999999) * |sweep| / 2 pi) segments of equal angle, each with arms of
// (4/3) tan(step / 4) along the ellipse tangents.
void PathData::arc_to(double px, double py, double rx, double ry, double x_rot_deg, bool large, bool sweep, double x,
                      double y) {
    const double kPi = 3.14159265358979323846;
    if (px == x && py == y) return;  // an arc to its own start is omitted (F.6.2)
    rx = std::fabs(rx);
    ry = std::fabs(ry);
    if (!(rx > 1e-5) || !(ry > 1e-5)) {  // F.6.2: a degenerate radius is a straight line
        line_to(x, y);
        return;
    }
    const double phi = x_rot_deg * kPi / 180.0;
    const double cp = std::cos(phi), sp = std::sin(phi);
    const double hx = (px - x) * 0.5, hy = (py - y) * 0.5;
    const double x1p = cp * hx + sp * hy, y1p = -sp * hx + cp * hy;
    const double lam = (x1p * x1p) / (rx * rx) + (y1p * y1p) / (ry * ry);
    if (lam > 1.0) {
        const double s = std::sqrt(lam);
        rx *= s;
        ry *= s;
    }
    const double rxs = rx * rx, rys = ry * ry;
    const double num = rxs * rys - rxs * y1p * y1p - rys * x1p * x1p;
    const double den = rxs * y1p * y1p + rys * x1p * x1p;
    double coef = den > 0 ? std::sqrt(std::fmax(num / den, 0.0)) : 0.0;
    if (large == sweep) coef = -coef;
    const double cxp = coef * rx * y1p / ry, cyp = -coef * ry * x1p / rx;
    const double cx = cp * cxp - sp * cyp + (px + x) * 0.5;
    const double cy = sp * cxp + cp * cyp + (py + y) * 0.5;
    auto angle = [](double ux, double uy, double vx, double vy) {
        return std::atan2(ux * vy - uy * vx, ux * vx + uy * vy);
    };
    const double ux = (x1p - cxp) / rx, uy = (y1p - cyp) / ry;
    const double vx = (-x1p - cxp) / rx, vy = (-y1p - cyp) / ry;
    const double start = angle(1.0, 0.0, ux, uy);
    double sweep_angle = angle(ux, uy, vx, vy);
    if (!sweep && sweep_angle > 0) sweep_angle -= 2.0 * kPi;
    if (sweep && sweep_angle < 0) sweep_angle += 2.0 * kPi;
    auto sample = [&](double a, double* ox, double* oy) {  // the ellipse at angle a, relative to its centre
        const double ex = rx * std::cos(a), ey = ry * std::sin(a);
        *ox = cp * ex - sp * ey;
        *oy = sp * ex + cp * ey;
    };
    const double tol = 0.1;
    const double scaled_err = std::fmax(rx, ry) / tol;
    const double n_err = std::fmax(std::pow(1.1163 * scaled_err, 1.0 / 6.0), 3.999999);
    const double n = std::ceil(n_err * std::fabs(sweep_angle) * (1.0 / (2.0 * kPi)));
    if (!(n >= 1.0)) {
        line_to(x, y);
        return;
    }
    const double step = sweep_angle / n;
    const double sign = sweep_angle < 0 ? -1.0 : 1.0;
    const double arm = (4.0 / 3.0) * std::fabs(std::tan(step / 4.0)) * sign;
    double a0 = start, p0x, p0y;
    sample(a0, &p0x, &p0y);
    for (int k = 0; k < (int)n; k++) {
        const double a1 = a0 + step;
        double t0x, t0y, p3x, p3y, t1x, t1y;
        sample(a0 + kPi / 2.0, &t0x, &t0y);
        sample(a1, &p3x, &p3y);
        sample(a1 + kPi / 2.0, &t1x, &t1y);
        curve_to(cx + p0x + arm * t0x, cy + p0y + arm * t0y, cx + p3x - arm * t1x, cy + p3y - arm * t1y, cx + p3x,
                 cy + p3y);
        a0 = a1;
        p0x = p3x;
        p0y = p3y;
    }
}

// Path data (the `d` attribute) -> absolute segments, usvg's rules.
PathData parse_path_data(const char* s, size_t len) {
    PathData path;
    Cursor c{s, s + len};
    double prev_mx = 0, prev_my = 0;  // the subpath start
    double prev_x = 0, prev_y = 0;    // the current point
    double prev_tx = 0, prev_ty = 0;  // the last T's control point
    char prev_cmd = 0;                // the previous command letter, upper case
    double prev_x2 = 0, prev_y2 = 0;  // the previous C/S second control point (absolute)
    double prev_qx = 0, prev_qy = 0;  // the previous Q control point (absolute)
    char cmd = 0;
    bool first = true;
    c.skip_ws();
    while (!c.done()) {
        const char ch = c.peek();
        if (std::isalpha((unsigned char)ch)) {
            cmd = ch;
            c.p++;
        } else if (cmd == 0 || std::toupper((unsigned char)cmd) == 'Z') {
            break;  // numbers without a command, or after Z: a parse error
        } else if (cmd == 'M') {
            cmd = 'L';  // coordinates after a moveto are implicit linetos
        } else if (cmd == 'm') {
            cmd = 'l';
        }
        const char up = (char)std::toupper((unsigned char)cmd);
        const bool abs = cmd == up;
        if (first && up != 'M') break;  // path data must start with a moveto
        first = false;
        bool ok = true;
        double v[7];
        bool fl[2];
        switch (up) {
        case 'M': {
            ok = list_number(c, &v[0]) && list_number(c, &v[1]);
            if (!ok) break;
            double x = v[0], y = v[1];
            if (!abs) {
                // after a Close, a relative moveto is from the subpath start
                if (!path.segs.empty() && path.segs.back().kind == S_CLOSE) {
                    x += prev_mx;
                    y += prev_my;
                } else {
                    x += prev_x;
                    y += prev_y;
                }
            }
            path.move_to(x, y);
            break;
        }
        case 'L': {
            ok = list_number(c, &v[0]) && list_number(c, &v[1]);
            if (!ok) break;
            double x = v[0], y = v[1];
            if (!abs) {
                x += prev_x;
                y += prev_y;
            }
            path.line_to(x, y);
            break;
        }
        case 'H': {
            ok = list_number(c, &v[0]);
            if (!ok) break;
            double x = v[0];
            if (!abs) x += prev_x;
            path.line_to(x, prev_y);
            break;
        }
        case 'V': {
            ok = list_number(c, &v[0]);
            if (!ok) break;
            double y = v[0];
            if (!abs) y += prev_y;
            path.line_to(prev_x, y);
            break;
        }
        case 'C': {
            for (int k = 0; k < 6 && ok; k++) ok = list_number(c, &v[k]);
            if (!ok) break;
            if (!abs)
                for (int k = 0; k < 6; k++) v[k] += (k & 1) ? prev_y : prev_x;
            path.curve_to(v[0], v[1], v[2], v[3], v[4], v[5]);
            prev_x2 = v[2];
            prev_y2 = v[3];
            break;
        }
        case 'S': {
            for (int k = 0; k < 4 && ok; k++) ok = list_number(c, &v[k]);
            if (!ok) break;
            if (!abs)
                for (int k = 0; k < 4; k++) v[k] += (k & 1) ? prev_y : prev_x;
            double x1 = prev_x, y1 = prev_y;
            if (prev_cmd == 'C' || prev_cmd == 'S') {
                x1 = prev_x * 2.0 - prev_x2;
                y1 = prev_y * 2.0 - prev_y2;
            }
            path.curve_to(x1, y1, v[0], v[1], v[2], v[3]);
            prev_x2 = v[0];
            prev_y2 = v[1];
            break;
        }
        case 'Q': {
            for (int k = 0; k < 4 && ok; k++) ok = list_number(c, &v[k]);
            if (!ok) break;
            if (!abs)
                for (int k = 0; k < 4; k++) v[k] += (k & 1) ? prev_y : prev_x;
            path.quad_to(prev_x, prev_y, v[0], v[1], v[2], v[3]);
            prev_qx = v[0];
            prev_qy = v[1];
            break;
        }
        case 'T': {
            ok = list_number(c, &v[0]) && list_number(c, &v[1]);
            if (!ok) break;
            if (!abs) {
                v[0] += prev_x;
                v[1] += prev_y;
            }
            double x1 = prev_x, y1 = prev_y;
            if (prev_cmd == 'Q') {
                x1 = prev_x * 2.0 - prev_qx;
                y1 = prev_y * 2.0 - prev_qy;
            } else if (prev_cmd == 'T') {
                x1 = prev_x * 2.0 - prev_tx;
                y1 = prev_y * 2.0 - prev_ty;
            }
            prev_tx = x1;
            prev_ty = y1;
            path.quad_to(prev_x, prev_y, x1, y1, v[0], v[1]);
            break;
        }
        case 'A': {
            ok = list_number(c, &v[0]) && list_number(c, &v[1]) && list_number(c, &v[2]) && list_flag(c, &fl[0]) &&
                 list_flag(c, &fl[1]) && list_number(c, &v[3]) && list_number(c, &v[4]);
            if (!ok) break;
            double x = v[3], y = v[4];
            if (!abs) {
                x += prev_x;
                y += prev_y;
            }
            path.arc_to(prev_x, prev_y, v[0], v[1], v[2], fl[0], fl[1], x, y);
            break;
        }
        case 'Z': {
            if (path.segs.empty() || path.segs.back().kind != S_CLOSE) path.close();
            c.skip_ws_comma();
            break;
        }
        default:
            ok = false;
            break;
        }
        if (!ok) break;
        prev_cmd = up;
        // the current point after the command
        if (!path.segs.empty()) {
            const Segment& sg = path.segs.back();
            if (sg.kind == S_MOVE) {
                prev_x = prev_mx = sg.x;
                prev_y = prev_my = sg.y;
            } else if (sg.kind == S_CLOSE) {
                prev_x = prev_mx;
                prev_y = prev_my;
            } else {
                prev_x = sg.x;
                prev_y = sg.y;
            }
        }
        c.skip_ws();
    }
    return path;
}

// ------------------------------------------------------------ transforms

bool parse_transform(const std::string& s, Affine* out) {
    const double kPi = 3.14159265358979323846;
    Cursor c{s.data(), s.data() + s.size()};
    Affine ts;
    c.skip_ws();
    while (!c.done()) {
        const char* name = c.p;
        while (!c.done() && std::isalpha((unsigned char)c.peek())) c.p++;
        const std::string fn(name, c.p);
        c.skip_ws();
        if (fn.empty() || c.peek() != '(') return false;
        c.p++;
        double v[6];
        int n = 0;
        c.skip_ws();
        while (!c.done() && c.peek() != ')' && n < 6) {
            if (!list_number(c, &v[n])) return false;
            n++;
        }
        c.skip_ws();
        if (c.peek() != ')') return false;
        c.p++;
        Affine t;
        if (fn == "matrix" && n == 6) {
            t = Affine{v[0], v[1], v[2], v[3], v[4], v[5]};
        } else if (fn == "translate" && (n == 1 || n == 2)) {
            t.e = v[0];
            t.f = n == 2 ? v[1] : 0.0;
        } else if (fn == "scale" && (n == 1 || n == 2)) {
            t.a = v[0];
            t.d = n == 2 ? v[1] : v[0];
        } else if (fn == "rotate" && (n == 1 || n == 3)) {
            const double r = v[0] * kPi / 180.0, ca = std::cos(r), sa = std::sin(r);
            Affine rot{ca, sa, -sa, ca, 0, 0};
            if (n == 3) {
                t = mul(mul(Affine{1, 0, 0, 1, v[1], v[2]}, rot), Affine{1, 0, 0, 1, -v[1], -v[2]});
            } else {
                t = rot;
            }
        } else if (fn == "skewX" && n == 1) {
            t.c = std::tan(v[0] * kPi / 180.0);
        } else if (fn == "skewY" && n == 1) {
            t.b = std::tan(v[0] * kPi / 180.0);
        } else {
            return false;
        }
        ts = mul(ts, t);
        c.skip_ws_comma();
    }
    *out = ts;
    return true;
}

// ------------------------------------------------------------ paint

struct Named {
    const char* name;
    uint32_t rgb;
};
// CSS Color Module level 3 extended colour keywords.
const Named kNamed[] = {
    {"aliceblue", 0xf0f8ff}, {"antiquewhite", 0xfaebd7}, {"aqua", 0x00ffff}, {"aquamarine", 0x7fffd4},
    {"azure", 0xf0ffff}, {"beige", 0xf5f5dc}, {"bisque", 0xffe4c4}, {"black", 0x000000},
    {"blanchedalmond", 0xffebcd}, {"blue", 0x0000ff}, {"blueviolet", 0x8a2be2}, {"brown", 0xa52a2a},
    {"burlywood", 0xdeb887}, {"cadetblue", 0x5f9ea0}, {"chartreuse", 0x7fff00}, {"chocolate", 0xd2691e},
    {"coral", 0xff7f50}, {"cornflowerblue", 0x6495ed}, {"cornsilk", 0xfff8dc}, {"crimson", 0xdc143c},
    {"cyan", 0x00ffff}, {"darkblue", 0x00008b}, {"darkcyan", 0x008b8b}, {"darkgoldenrod", 0xb8860b},
    {"darkgray", 0xa9a9a9}, {"darkgreen", 0x006400}, {"darkgrey", 0xa9a9a9}, {"darkkhaki", 0xbdb76b},
    {"darkmagenta", 0x8b008b}, {"darkolivegreen", 0x556b2f}, {"darkorange", 0xff8c00}, {"darkorchid", 0x9932cc},
    {"darkred", 0x8b0000}, {"darksalmon", 0xe9967a}, {"darkseagreen", 0x8fbc8f}, {"darkslateblue", 0x483d8b},
    {"darkslategray", 0x2f4f4f}, {"darkslategrey", 0x2f4f4f}, {"darkturquoise", 0x00ced1},
    {"darkviolet", 0x9400d3}, {"deeppink", 0xff1493}, {"deepskyblue", 0x00bfff}, {"dimgray", 0x696969},
    {"dimgrey", 0x696969}, {"dodgerblue", 0x1e90ff}, {"firebrick", 0xb22222}, {"floralwhite", 0xfffaf0},
    {"forestgreen", 0x228b22}, {"fuchsia", 0xff00ff}, {"gainsboro", 0xdcdcdc}, {"ghostwhite", 0xf8f8ff},
    {"gold", 0xffd700}, {"goldenrod", 0xdaa520}, {"gray", 0x808080}, {"grey", 0x808080}, {"green", 0x008000},
    {"greenyellow", 0xadff2f}, {"honeydew", 0xf0fff0}, {"hotpink", 0xff69b4}, {"indianred", 0xcd5c5c},
    {"indigo", 0x4b0082}, {"ivory", 0xfffff0}, {"khaki", 0xf0e68c}, {"lavender", 0xe6e6fa},
    {"lavenderblush", 0xfff0f5}, {"lawngreen", 0x7cfc00}, {"lemonchiffon", 0xfffacd}, {"lightblue", 0xadd8e6},
    {"lightcoral", 0xf08080}, {"lightcyan", 0xe0ffff}, {"lightgoldenrodyellow", 0xfafad2},
    {"lightgray", 0xd3d3d3}, {"lightgreen", 0x90ee90}, {"lightgrey", 0xd3d3d3}, {"lightpink", 0xffb6c1},
    {"lightsalmon", 0xffa07a}, {"lightseagreen", 0x20b2aa}, {"lightskyblue", 0x87cefa},
    {"lightslategray", 0x778899}, {"lightslategrey", 0x778899}, {"lightsteelblue", 0xb0c4de},
    {"lightyellow", 0xffffe0}, {"lime", 0x00ff00}, {"limegreen", 0x32cd32}, {"linen", 0xfaf0e6},
    {"magenta", 0xff00ff}, {"maroon", 0x800000}, {"mediumaquamarine", 0x66cdaa}, {"mediumblue", 0x0000cd},
    {"mediumorchid", 0xba55d3}, {"mediumpurple", 0x9370db}, {"mediumseagreen", 0x3cb371},
    {"mediumslateblue", 0x7b68ee}, {"mediumspringgreen", 0x00fa9a}, {"mediumturquoise", 0x48d1cc},
    {"mediumvioletred", 0xc71585}, {"midnightblue", 0x191970}, {"mintcream", 0xf5fffa},
    {"mistyrose", 0xffe4e1}, {"moccasin", 0xffe4b5}, {"navajowhite", 0xffdead}, {"navy", 0x000080},
    {"oldlace", 0xfdf5e6}, {"olive", 0x808000}, {"olivedrab", 0x6b8e23}, {"orange", 0xffa500},
    {"orangered", 0xff4500}, {"orchid", 0xda70d6}, {"palegoldenrod", 0xeee8aa}, {"palegreen", 0x98fb98},
    {"paleturquoise", 0xafeeee}, {"palevioletred", 0xdb7093}, {"papayawhip", 0xffefd5},
    {"peachpuff", 0xffdab9}, {"peru", 0xcd853f}, {"pink", 0xffc0cb}, {"plum", 0xdda0dd},
    {"powderblue", 0xb0e0e6}, {"purple", 0x800080}, {"red", 0xff0000}, {"rosybrown", 0xbc8f8f},
    {"royalblue", 0x4169e1}, {"saddlebrown", 0x8b4513}, {"salmon", 0xfa8072}, {"sandybrown", 0xf4a460},
    {"seagreen", 0x2e8b57}, {"seashell", 0xfff5ee}, {"sienna", 0xa0522d}, {"silver", 0xc0c0c0},
    {"skyblue", 0x87ceeb}, {"slateblue", 0x6a5acd}, {"slategray", 0x708090}, {"slategrey", 0x708090},
    {"snow", 0xfffafa}, {"springgreen", 0x00ff7f}, {"steelblue", 0x4682b4}, {"tan", 0xd2b48c},
    {"teal", 0x008080}, {"thistle", 0xd8bfd8}, {"tomato", 0xff6347}, {"turquoise", 0x40e0d0},
    {"violet", 0xee82ee}, {"wheat", 0xf5deb3}, {"white", 0xffffff}, {"whitesmoke", 0xf5f5f5},
    {"yellow", 0xffff00}, {"yellowgreen", 0x9acd32},
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && is_ws(s[a])) a++;
    while (b > a && is_ws(s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string lower(std::string s) {
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

int hexval(char ch) {
    if (ch >= '0' && ch <= '9') return ch - '0';
    ch = (char)std::tolower((unsigned char)ch);
    if (ch >= 'a' && ch <= 'f') return ch - 'a' + 10;
    return -1;
}

// A colour value: #rgb, #rrggbb, rgb(r, g, b) (integers or percentages) or a
// keyword.  Returns false for anything else.
bool parse_color(const std::string& raw, uint32_t* rgb) {
    const std::string s = trim(raw);
    if (!s.empty() && s[0] == '#') {
        const std::string h = s.substr(1);
        int v[6];
        for (size_t i = 0; i < h.size() && i < 6; i++) v[i] = hexval(h[i]);
        if (h.size() == 3 && v[0] >= 0 && v[1] >= 0 && v[2] >= 0) {
            *rgb = (uint32_t)((v[0] * 17) << 16 | (v[1] * 17) << 8 | (v[2] * 17));
            return true;
        }
        if (h.size() == 6) {
            for (int i = 0; i < 6; i++)
                if (v[i] < 0) return false;
            *rgb = (uint32_t)((v[0] * 16 + v[1]) << 16 | (v[2] * 16 + v[3]) << 8 | (v[4] * 16 + v[5]));
            return true;
        }
        return false;
    }
    const std::string l = lower(s);
    if (l.compare(0, 4, "rgb(") == 0 && l.back() == ')') {
        Cursor c{l.data() + 4, l.data() + l.size() - 1};
        int ch[3];
        for (int k = 0; k < 3; k++) {
            double v;
            c.skip_ws();
            if (!parse_number(c, &v)) return false;
            if (c.peek() == '%') {
                c.p++;
                v = v * 255.0 / 100.0;
            }
            v = std::round(v);
            ch[k] = v < 0 ? 0 : (v > 255 ? 255 : (int)v);
            c.skip_ws_comma();
        }
        *rgb = (uint32_t)(ch[0] << 16 | ch[1] << 8 | ch[2]);
        return true;
    }
    for (const Named& n : kNamed) {
        if (l == n.name) {
            *rgb = n.rgb;
            return true;
        }
    }
    return false;
}

// A length in user units: a number with an optional absolute unit.
bool parse_length(const std::string& raw, double* out) {
    const std::string s = trim(raw);
    Cursor c{s.data(), s.data() + s.size()};
    double v;
    if (!parse_number(c, &v)) return false;
    const std::string unit(c.p, c.end);
    double k = 1.0;
    if (unit.empty() || unit == "px") k = 1.0;
    else if (unit == "pt") k = 4.0 / 3.0;
    else if (unit == "pc") k = 16.0;
    else if (unit == "mm") k = 3.7795275590551181;
    else if (unit == "cm") k = 37.795275590551181;
    else if (unit == "in") k = 96.0;
    else return false;
    *out = v * k;
    return true;
}

// An opacity: a number (or a percentage), clamped to [0, 1].
bool parse_opacity(const std::string& raw, double* out) {
    const std::string s = trim(raw);
    Cursor c{s.data(), s.data() + s.size()};
    double v;
    if (!parse_number(c, &v)) return false;
    if (c.peek() == '%') v /= 100.0;
    *out = v < 0 ? 0 : (v > 1 ? 1 : v);
    return true;
}

// ------------------------------------------------------------ XML

struct Element {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<Element> children;
    const std::string* attr(const char* k) const {
        for (const auto& a : attrs)
            if (a.first == k) return &a.second;
        return nullptr;
    }
};

std::string decode_entities(const char* s, const char* e) {
    std::string out;
    out.reserve((size_t)(e - s));
    while (s < e) {
        if (*s == '&') {
            const char* semi = (const char*)std::memchr(s, ';', (size_t)(e - s));
            if (semi) {
                const std::string ent(s + 1, semi);
                bool done = true;
                if (ent == "amp") out += '&';
                else if (ent == "lt") out += '<';
                else if (ent == "gt") out += '>';
                else if (ent == "quot") out += '"';
                else if (ent == "apos") out += '\'';
                else if (!ent.empty() && ent[0] == '#') {
                    long code = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')
                                    ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                    : std::strtol(ent.c_str() + 1, nullptr, 10);
                    if (code > 0 && code < 128) out += (char)code;
                    else out += ' ';  // outside ASCII: not part of any value read here
                } else {
                    done = false;
                }
                if (done) {
                    s = semi + 1;
                    continue;
                }
            }
        }
        out += *s++;
    }
    return out;
}

// The name without its namespace prefix.
std::string local_name(const std::string& q) {
    const size_t c = q.find(':');
    return c == std::string::npos ? q : q.substr(c + 1);
}

struct XmlParser {
    const char* p;
    const char* end;
    bool error = false;

    bool starts(const char* lit) const {
        const size_t n = std::strlen(lit);
        return (size_t)(end - p) >= n && std::memcmp(p, lit, n) == 0;
    }
    bool skip_past(const char* lit) {
        const size_t n = std::strlen(lit);
        while (p + n <= end) {
            if (std::memcmp(p, lit, n) == 0) {
                p += n;
                return true;
            }
            p++;
        }
        p = end;
        return false;
    }
    // Skips text, comments, processing instructions, DOCTYPE and CDATA up to
    // the next tag; returns false at the end of input.
    bool next_tag() {
        for (;;) {
            while (p < end && *p != '<') p++;
            if (p >= end) return false;
            if (starts("<!--")) {
                if (!skip_past("-->")) return false;
            } else if (starts("<?")) {
                if (!skip_past("?>")) return false;
            } else if (starts("<![CDATA[")) {
                if (!skip_past("]]>")) return false;
            } else if (starts("<!")) {
                int depth = 0;  // DOCTYPE with an optional internal subset
                while (p < end) {
                    if (*p == '[') depth++;
                    else if (*p == ']') depth--;
                    else if (*p == '>' && depth <= 0) break;
                    p++;
                }
                if (p >= end) return false;
                p++;
            } else {
                return true;
            }
        }
    }
    std::string name() {
        const char* s = p;
        while (p < end && !is_ws(*p) && *p != '>' && *p != '/' && *p != '=') p++;
        return std::string(s, p);
    }
    void ws() {
        while (p < end && is_ws(*p)) p++;
    }
    // Parses the element whose '<' is at p, with its subtree.
    bool element(Element& el, int depth) {
        if (depth > 256) return false;
        p++;  // '<'
        el.name = local_name(name());
        if (el.name.empty()) return false;
        for (;;) {
            ws();
            if (p >= end) return false;
            if (*p == '/') {
                if (p + 1 < end && p[1] == '>') {
                    p += 2;
                    return true;
                }
                return false;
            }
            if (*p == '>') {
                p++;
                break;
            }
            std::string k = name();
            ws();
            if (p >= end || *p != '=') return false;
            p++;
            ws();
            if (p >= end || (*p != '"' && *p != '\'')) return false;
            const char q = *p++;
            const char* vs = p;
            while (p < end && *p != q) p++;
            if (p >= end) return false;
            el.attrs.emplace_back(local_name(k), decode_entities(vs, p));
            p++;
        }
        // children up to the end tag
        for (;;) {
            if (!next_tag()) return false;
            if (starts("</")) {
                p += 2;
                const std::string n = local_name(name());
                ws();
                if (p >= end || *p != '>' || n != el.name) return false;
                p++;
                return true;
            }
            el.children.emplace_back();
            if (!element(el.children.back(), depth + 1)) return false;
        }
    }
};

// ------------------------------------------------------------ the tree walk

struct Paint {
    int kind = 0;  // 0 none, 1 colour, 2 other (a url(#...) server: not drawn by the demo)
    uint32_t rgb = 0;
};

struct Style {
    Paint fill{1, 0};  // black
    Paint stroke;      // none
    double fill_opacity = 1.0, stroke_opacity = 1.0, stroke_width = 1.0;
    uint32_t color = 0;  // `color`, for currentColor
};

// Presentation attributes of an element; `style` declarations override them.
std::vector<std::pair<std::string, std::string>> properties(const Element& el) {
    std::vector<std::pair<std::string, std::string>> props;
    for (const auto& a : el.attrs)
        if (a.first != "style") props.push_back(a);
    if (const std::string* st = el.attr("style")) {
        size_t i = 0;
        while (i < st->size()) {
            size_t semi = st->find(';', i);
            if (semi == std::string::npos) semi = st->size();
            const std::string decl = st->substr(i, semi - i);
            const size_t colon = decl.find(':');
            if (colon != std::string::npos) props.emplace_back(trim(decl.substr(0, colon)), trim(decl.substr(colon + 1)));
            i = semi + 1;
        }
    }
    return props;
}

bool parse_paint(const std::string& raw, const Style& st, const Paint& inherited, Paint* out) {
    const std::string s = trim(raw);
    if (s == "none") {
        *out = Paint{0, 0};
        return true;
    }
    if (s == "inherit") {
        *out = inherited;
        return true;
    }
    if (s == "currentColor") {
        *out = Paint{1, st.color};
        return true;
    }
    if (s.compare(0, 4, "url(") == 0) {
        *out = Paint{2, 0};
        return true;
    }
    uint32_t rgb;
    if (parse_color(s, &rgb)) {
        *out = Paint{1, rgb};
        return true;
    }
    return false;  // an invalid value leaves the inherited paint
}

struct Doc {
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    std::vector<int64_t> verb_off{0}, point_off{0};
    std::vector<ochre_svg_draw> draws;
    double view_box[4] = {0, 0, 0, 0};
    double size[2] = {0, 0};
    int64_t skipped = 0;  // elements not rendered (unsupported or not graphics)
};

uint8_t opacity_u8(double o) { return (uint8_t)std::ceil(o * 255.0); }

void add_path(Doc& doc, const PathData& pd, const Style& st, const Affine& ts) {
    if (pd.segs.size() < 2) return;
    const int64_t index = (int64_t)doc.verb_off.size() - 1;
    for (const Segment& s : pd.segs) {
        switch (s.kind) {
        case S_MOVE:
            doc.verbs.push_back(OCHRE_VERB_MOVE);
            doc.points.push_back((float)s.x);
            doc.points.push_back((float)s.y);
            break;
        case S_LINE:
            doc.verbs.push_back(OCHRE_VERB_LINE);
            doc.points.push_back((float)s.x);
            doc.points.push_back((float)s.y);
            break;
        case S_CURVE:
            doc.verbs.push_back(OCHRE_VERB_CUBIC);
            for (double v : {s.x1, s.y1, s.x2, s.y2, s.x, s.y}) doc.points.push_back((float)v);
            break;
        case S_CLOSE:
            doc.verbs.push_back(OCHRE_VERB_CLOSE);
            break;
        }
    }
    doc.verb_off.push_back((int64_t)doc.verbs.size());
    doc.point_off.push_back((int64_t)doc.points.size() / 2);
    ochre_svg_draw d;
    std::memset(&d, 0, sizeof d);
    d.path = index;
    // svg.rs:107-111: Mat2x2::new(a, c, b, d) (row-major) and offset (e, f)
    d.transform.m[0] = (float)ts.a;
    d.transform.m[1] = (float)ts.c;
    d.transform.m[2] = (float)ts.b;
    d.transform.m[3] = (float)ts.d;
    d.transform.offset[0] = (float)ts.e;
    d.transform.offset[1] = (float)ts.f;
    if (st.fill.kind == 1) {
        d.stroke = 0;
        d.width = 0.0f;
        d.color[0] = (uint8_t)(st.fill.rgb >> 16);
        d.color[1] = (uint8_t)(st.fill.rgb >> 8);
        d.color[2] = (uint8_t)st.fill.rgb;
        d.color[3] = opacity_u8(st.fill_opacity);
        doc.draws.push_back(d);
    }
    if (st.stroke.kind == 1 && st.stroke_width > 0.0) {
        d.stroke = 1;
        d.width = (float)st.stroke_width;
        d.color[0] = (uint8_t)(st.stroke.rgb >> 16);
        d.color[1] = (uint8_t)(st.stroke.rgb >> 8);
        d.color[2] = (uint8_t)st.stroke.rgb;
        d.color[3] = opacity_u8(st.stroke_opacity);
        doc.draws.push_back(d);
    }
}

double num_attr(const Element& el, const char* k, double dflt) {
    const std::string* v = el.attr(k);
    double out;
    return v && parse_length(*v, &out) ? out : dflt;
}

// A list of numbers (polyline / polygon points).
std::vector<double> number_list(const std::string& s) {
    std::vector<double> out;
    Cursor c{s.data(), s.data() + s.size()};
    double v;
    while (list_number(c, &v)) out.push_back(v);
    return out;
}

// Shape elements -> path data (usvg's shapes module).
bool shape_path(const Element& el, PathData* pd) {
    const std::string& n = el.name;
    if (n == "path") {
        const std::string* d = el.attr("d");
        if (!d) return false;
        *pd = parse_path_data(d->data(), d->size());
        return true;
    }
    if (n == "rect") {
        const double x = num_attr(el, "x", 0), y = num_attr(el, "y", 0);
        const double w = num_attr(el, "width", 0), h = num_attr(el, "height", 0);
        if (!(w > 0) || !(h > 0)) return false;
        const std::string* rxs = el.attr("rx");
        const std::string* rys = el.attr("ry");
        double rx = num_attr(el, "rx", 0), ry = num_attr(el, "ry", 0);
        if (rxs && !rys) ry = rx;
        if (rys && !rxs) rx = ry;
        rx = std::fmin(rx, w / 2);
        ry = std::fmin(ry, h / 2);
        if (rx <= 0 || ry <= 0) {
            pd->move_to(x, y);
            pd->line_to(x + w, y);
            pd->line_to(x + w, y + h);
            pd->line_to(x, y + h);
            pd->close();
        } else {
            pd->move_to(x + rx, y);
            pd->line_to(x + w - rx, y);
            pd->arc_to(x + w - rx, y, rx, ry, 0, false, true, x + w, y + ry);
            pd->line_to(x + w, y + h - ry);
            pd->arc_to(x + w, y + h - ry, rx, ry, 0, false, true, x + w - rx, y + h);
            pd->line_to(x + rx, y + h);
            pd->arc_to(x + rx, y + h, rx, ry, 0, false, true, x, y + h - ry);
            pd->line_to(x, y + ry);
            pd->arc_to(x, y + ry, rx, ry, 0, false, true, x + rx, y);
            pd->close();
        }
        return true;
    }
    if (n == "circle" || n == "ellipse") {
        const double cx = num_attr(el, "cx", 0), cy = num_attr(el, "cy", 0);
        double rx, ry;
        if (n == "circle") {
            rx = ry = num_attr(el, "r", 0);
        } else {
            rx = num_attr(el, "rx", 0);
            ry = num_attr(el, "ry", 0);
        }
        if (!(rx > 0) || !(ry > 0)) return false;
        pd->move_to(cx + rx, cy);
        pd->arc_to(cx + rx, cy, rx, ry, 0, false, true, cx, cy + ry);
        pd->arc_to(cx, cy + ry, rx, ry, 0, false, true, cx - rx, cy);
        pd->arc_to(cx - rx, cy, rx, ry, 0, false, true, cx, cy - ry);
        pd->arc_to(cx, cy - ry, rx, ry, 0, false, true, cx + rx, cy);
        pd->close();
        return true;
    }
    if (n == "line") {
        pd->move_to(num_attr(el, "x1", 0), num_attr(el, "y1", 0));
        pd->line_to(num_attr(el, "x2", 0), num_attr(el, "y2", 0));
        return true;
    }
    if (n == "polyline" || n == "polygon") {
        const std::string* pts = el.attr("points");
        if (!pts) return false;
        const std::vector<double> v = number_list(*pts);
        if (v.size() < 4) return false;
        pd->move_to(v[0], v[1]);
        for (size_t i = 2; i + 1 < v.size(); i += 2) pd->line_to(v[i], v[i + 1]);
        if (n == "polygon") pd->close();
        return true;
    }
    return false;
}

bool is_container(const std::string& n) { return n == "g" || n == "svg" || n == "a" || n == "switch"; }
bool is_shape(const std::string& n) {
    return n == "path" || n == "rect" || n == "circle" || n == "ellipse" || n == "line" || n == "polyline" ||
           n == "polygon";
}

// Walks an element: `st` is the inherited style, `ts` the transform since the
// nearest kept group.
void walk(Doc& doc, const Element& el, Style st, Affine ts) {
    if (!is_container(el.name) && !is_shape(el.name)) {
        doc.skipped++;
        return;
    }
    bool kept = false;  // opacity, clip-path, mask or filter keep a group in usvg's tree
    Affine own;
    bool display_none = false;
    for (const auto& kv : properties(el)) {
        const std::string& k = kv.first;
        const std::string& v = kv.second;
        if (k == "fill") parse_paint(v, st, st.fill, &st.fill);
        else if (k == "stroke") parse_paint(v, st, st.stroke, &st.stroke);
        else if (k == "fill-opacity") parse_opacity(v, &st.fill_opacity);
        else if (k == "stroke-opacity") parse_opacity(v, &st.stroke_opacity);
        else if (k == "stroke-width") parse_length(v, &st.stroke_width);
        else if (k == "color") parse_color(v, &st.color);
        else if (k == "display" && trim(v) == "none") display_none = true;
        else if (k == "opacity") {
            double o = 1.0;
            if (parse_opacity(v, &o) && o != 1.0) kept = true;
        } else if ((k == "clip-path" || k == "mask" || k == "filter") && trim(v) != "none") {
            kept = true;
        }
    }
    if (display_none) {
        doc.skipped++;
        return;
    }
    if (const std::string* t = el.attr("transform")) {
        if (!parse_transform(*t, &own)) own = Affine{};
    }
    // the root <svg>'s own attributes (viewBox, size) are not a transform
    const Affine here = kept ? Affine{} : mul(ts, own);
    if (is_shape(el.name)) {
        PathData pd;
        if (shape_path(el, &pd)) add_path(doc, pd, st, here);
        else doc.skipped++;
        return;
    }
    for (const Element& ch : el.children) walk(doc, ch, st, here);
}

}  // namespace

struct ochre_svg_document {
    Doc doc;
};

extern "C" {

int ochre_svg_parse(const char* text, size_t len, ochre_svg_document** out) {
    if (!out) return OCHRE_ERR_INVALID_ARG;
    *out = nullptr;
    if (!text) return OCHRE_ERR_INVALID_ARG;
    try {
        XmlParser x{text, text + len};
        if (!x.next_tag()) return OCHRE_ERR_INVALID_ARG;
        Element root;
        if (!x.element(root, 0) || root.name != "svg") return OCHRE_ERR_INVALID_ARG;
        ochre_svg_document* d = new ochre_svg_document();
        if (const std::string* vb = root.attr("viewBox")) {
            const std::vector<double> v = number_list(*vb);
            if (v.size() == 4)
                for (int i = 0; i < 4; i++) d->doc.view_box[i] = v[i];
        }
        d->doc.size[0] = num_attr(root, "width", d->doc.view_box[2]);
        d->doc.size[1] = num_attr(root, "height", d->doc.view_box[3]);
        // the root <svg> element's own transform attribute is ignored (SVG 1.1)
        Element top = root;
        top.attrs.erase(std::remove_if(top.attrs.begin(), top.attrs.end(),
                                       [](const std::pair<std::string, std::string>& a) { return a.first == "transform"; }),
                        top.attrs.end());
        walk(d->doc, top, Style{}, Affine{});
        *out = d;
        return OCHRE_OK;
    } catch (const std::bad_alloc&) {
        return OCHRE_ERR_OUT_OF_MEMORY;
    }
}

int ochre_svg_info(const ochre_svg_document* d, ochre_svg_summary* info) {
    if (!d || !info) return OCHRE_ERR_INVALID_ARG;
    info->n_paths = (int64_t)d->doc.verb_off.size() - 1;
    info->n_verbs = (int64_t)d->doc.verbs.size();
    info->n_points = (int64_t)d->doc.points.size() / 2;
    info->n_draws = (int64_t)d->doc.draws.size();
    info->n_skipped = d->doc.skipped;
    for (int i = 0; i < 4; i++) info->view_box[i] = (float)d->doc.view_box[i];
    info->width = (float)d->doc.size[0];
    info->height = (float)d->doc.size[1];
    return OCHRE_OK;
}

const uint8_t* ochre_svg_verbs(const ochre_svg_document* d) { return d ? d->doc.verbs.data() : nullptr; }
const int64_t* ochre_svg_verb_offsets(const ochre_svg_document* d) { return d ? d->doc.verb_off.data() : nullptr; }
const float* ochre_svg_points(const ochre_svg_document* d) { return d ? d->doc.points.data() : nullptr; }
const int64_t* ochre_svg_point_offsets(const ochre_svg_document* d) { return d ? d->doc.point_off.data() : nullptr; }
const ochre_svg_draw* ochre_svg_draws(const ochre_svg_document* d) { return d ? d->doc.draws.data() : nullptr; }

void ochre_svg_free(ochre_svg_document* d) { delete d; }

}  // extern "C"
