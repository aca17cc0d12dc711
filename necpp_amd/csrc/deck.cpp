// deck.cpp -- the card-deck front end: NEC input text driven through the
// libnecpp calls, so that decks written for nec2++ run on the GPU path.
//
// The job follows nec2++'s main loop (src/nec2cpp.cpp:290-440): optional
// comment cards (CM ... CE), the geometry section up to GE
// (src/c_geometry.cpp:115-165), then data cards until EN or the end of the
// text.  Lines are read as src/misc.cpp:74-160 reads them: a line whose first
// character is '#' or a blank is skipped, a single quote starts a comment,
// the first two characters are the card's mnemonic (case-insensitive).
// Fields are separated by blanks, tabs or commas; data cards carry four
// integers and six reals (src/nec_card_parser.h:40-75), geometry cards two
// integers and seven reals (src/c_geometry.cpp:2600-2735); a missing field
// reads as zero.
//
// Geometry cards: GW (with its GC continuation when the radius is zero, a
// tapered wire), GM (move / rotate, with copies), GS (scales everything
// defined so far) and GE, each applied to the segments as it is read.  Data
// cards go to the libnecpp function of the same name; cards outside the GPU
// path report that as the API does.  The first failing card stops the deck
// with an error naming the card; results computed before it stay in the
// context.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/libnecpp.h"
#include "context.h"
#include "nec_math.h"

namespace necgpu {

namespace {

struct Card {
  std::string name;  // mnemonic, upper case
  std::string rest;  // the fields
  int line = 0;
};

bool is_sep(char c) { return c == ' ' || c == '\t' || c == ','; }

std::vector<std::string> tokens(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && is_sep(s[i])) ++i;
    size_t j = i;
    while (j < s.size() && !is_sep(s[j])) ++j;
    if (j > i) out.push_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

// The cards of the text in order (skipped lines and comments removed).
std::vector<Card> read_cards(const char* text) {
  std::vector<Card> cards;
  std::string s(text ? text : "");
  size_t pos = 0;
  int line = 0;
  while (pos <= s.size()) {
    size_t end = s.find_first_of("\r\n", pos);
    if (end == std::string::npos) end = s.size();
    std::string l = s.substr(pos, end - pos);
    ++line;
    pos = end + 1;
    if (end < s.size() && s[end] == '\r' && pos < s.size() && s[pos] == '\n') ++pos;
    if (l.empty() || l[0] == '#' || l[0] == ' ') continue;
    const size_t q = l.find('\'');
    if (q != std::string::npos) l = l.substr(0, q);
    if (l.empty()) continue;
    Card c;
    c.name = l.substr(0, std::min<size_t>(2, l.size()));
    for (char& ch : c.name) ch = (char)std::toupper((unsigned char)ch);
    c.rest = l.size() > 2 ? l.substr(2) : std::string();
    c.line = line;
    cards.push_back(c);
    if (end >= s.size()) break;
  }
  return cards;
}

std::string where(const Card& c) { return "deck line " + std::to_string(c.line) + " (" + c.name + ")"; }

// ni integers then nr reals; geometry cards refuse characters that cannot be
// part of a number.
void fields(const Card& c, int ni, int nr, bool strict, int* iv, double* rv) {
  for (int k = 0; k < ni; ++k) iv[k] = 0;
  for (int k = 0; k < nr; ++k) rv[k] = 0.0;
  const std::vector<std::string> t = tokens(c.rest);
  for (size_t k = 0; k < t.size() && (int)k < ni + nr; ++k) {
    const bool integer = (int)k < ni;
    if (strict)
      for (char ch : t[k])
        if (!(std::isdigit((unsigned char)ch) || ch == '+' || ch == '-' ||
              (!integer && (ch == '.' || ch == 'e' || ch == 'E'))))
          throw NecError("GEOMETRY DATA CARD \"" + c.name + "\" ERROR: NON-NUMERICAL CHARACTER '" +
                         std::string(1, ch) + "' IN " + (integer ? "INTEGER" : "FLOAT") + " FIELD (" + where(c) + ")");
    if (integer)
      iv[k] = std::atoi(t[k].c_str());
    else
      rv[k - ni] = std::atof(t[k].c_str());
  }
}

void check(long rc, const Card& c) {
  if (rc != 0) throw NecError(where(c) + ": " + nec_error_message());
}

}  // namespace

void run_deck(Context* ctx, const char* text) {
  if (!ctx) throw NecError("deck: no context");
  const std::vector<Card> cards = read_cards(text);
  size_t k = 0;
  // comment cards: CM ... CE (src/nec2cpp.cpp:318-344)
  if (k < cards.size() && (cards[k].name == "CM" || cards[k].name == "CE")) {
    while (k < cards.size() && cards[k].name == "CM") ++k;
    if (k >= cards.size() || cards[k].name != "CE") throw NecError("ERROR: INCORRECT LABEL FOR A COMMENT CARD");
    ++k;
  }
  // geometry section: each card acts on the segments defined so far, as
  // c_geometry::parse_geometry does (src/c_geometry.cpp:115-165)
  bool complete = false;
  int iv[4];
  double rv[7];
  while (k < cards.size() && !complete) {
    const Card& c = cards[k++];
    if (c.name == "GW") {
      fields(c, 2, 7, true, iv, rv);
      const int tag = iv[0], ns = iv[1];
      double rad = rv[6], rdel = 1.0, rrad = 1.0;
      if (rad == 0.0) {
        // tapered wire: GC gives the segment length ratio and the end radii
        // (src/c_geometry.cpp:189-234)
        if (k >= cards.size() || cards[k].name != "GC") throw NecError("GEOMETRY DATA CARD ERROR (" + where(c) + ")");
        const Card& g = cards[k++];
        double gv[7];
        fields(g, 2, 7, true, iv, gv);
        if (gv[1] == 0.0 || gv[2] == 0.0) throw NecError("GEOMETRY DATA CARD ERROR (" + where(g) + ")");
        rdel = gv[0];
        rad = gv[1];
        rrad = std::pow(gv[2] / gv[1], 1.0 / (ns - 1.0));
      }
      check(nec_wire(ctx, tag, ns, rv[0], rv[1], rv[2], rv[3], rv[4], rv[5], rad, rdel, rrad), c);
    } else if (c.name == "GS") {
      // scale every segment defined so far, coordinates and radii
      // (c_geometry::scale, src/c_geometry.cpp:1308-1318)
      fields(c, 2, 7, true, iv, rv);
      ctx->scale(rv[0]);
    } else if (c.name == "GE") {
      fields(c, 2, 7, true, iv, rv);
      check(nec_geometry_complete(ctx, iv[0]), c);
      complete = true;
    } else if (c.name == "GM") {
      // move / rotate, optionally as copies (src/c_geometry.cpp:256-268):
      // I1 the tag increment, I2 the copies, F1-F3 the rotations (degrees),
      // F4-F6 the translation, F7 the first tag moved; the same call as
      // nec_gm_card, so decks and API programs build the same segments
      fields(c, 2, 7, true, iv, rv);
      check(nec_gm_card(ctx, iv[0], iv[1], rv[0], rv[1], rv[2], rv[3], rv[4], rv[5], static_cast<int>(rv[6] + .5)),
            c);
    } else if (c.name == "GA" || c.name == "GH" || c.name == "GR" || c.name == "GX" ||
               c.name == "SP" || c.name == "SM" || c.name == "SC" || c.name == "GF") {
      throw NecError(where(c) + ": " + c.name + ": not supported by the GPU path");
    } else {
      throw NecError("GEOMETRY DATA CARD ERROR: unknown card (" + where(c) + ")");
    }
  }
  if (!complete) throw NecError("deck: no GE card ends the geometry");
  // data cards (src/nec2cpp.cpp:360-440, handlers :557-625)
  double f[6];
  while (k < cards.size()) {
    const Card& c = cards[k++];
    const std::string& m = c.name;
    if (m == "EN") return;
    fields(c, 4, 6, false, iv, f);
    if (m == "FR") check(nec_fr_card(ctx, iv[0], iv[1], f[0], f[1]), c);
    else if (m == "EX") check(nec_ex_card(ctx, iv[0], iv[1], iv[2], iv[3], f[0], f[1], f[2], f[3], f[4], f[5]), c);
    else if (m == "LD") check(nec_ld_card(ctx, iv[0], iv[1], iv[2], iv[3], f[0], f[1], f[2]), c);
    else if (m == "GN") check(nec_gn_card(ctx, iv[0], iv[1], f[0], f[1], f[2], f[3], f[4], f[5]), c);
    else if (m == "EK") check(nec_ek_card(ctx, iv[0]), c);
    else if (m == "KH") check(nec_kh_card(ctx, f[0]), c);
    else if (m == "XQ") check(nec_xq_card(ctx, iv[0]), c);
    else if (m == "RP") {
      // I4 packs the output format, normalisation, gain type and averaging
      // as the digits X N D A
      const int xnda = iv[3];
      check(nec_rp_card(ctx, iv[0], iv[1], iv[2], xnda / 1000, (xnda / 100) % 10, (xnda / 10) % 10, xnda % 10, f[0],
                        f[1], f[2], f[3], f[4], f[5]),
            c);
    } else if (m == "PT") check(nec_pt_card(ctx, iv[0], iv[1], iv[2], iv[3]), c);
    else if (m == "PQ") check(nec_pq_card(ctx, iv[0], iv[1], iv[2], iv[3]), c);
    else if (m == "TL") check(nec_tl_card(ctx, iv[0], iv[1], iv[2], iv[3], f[0], f[1], f[2], f[3], f[4], f[5]), c);
    else if (m == "NT") check(nec_nt_card(ctx, iv[0], iv[1], iv[2], iv[3], f[0], f[1], f[2], f[3], f[4], f[5]), c);
    else if (m == "NE") check(nec_ne_card(ctx, iv[0], iv[1], iv[2], iv[3], f[0], f[1], f[2], f[3], f[4], f[5]), c);
    else if (m == "NH") check(nec_nh_card(ctx, iv[0], iv[1], iv[2], iv[3], f[0], f[1], f[2], f[3], f[4], f[5]), c);
    else if (m == "CP") check(nec_cp_card(ctx, iv[0], iv[1], iv[2], iv[3]), c);
    else if (m == "GD") check(nec_gd_card(ctx, f[0], f[1], f[2], f[3]), c);
    else if (m == "MP") check(nec_medium_parameters(ctx, f[0], f[1]), c);
    else if (m == "PL" || m == "WG" || m == "NX")
      throw NecError(where(c) + ": " + m + ": not supported by the GPU path");
    else throw NecError("FAULTY DATA CARD LABEL AFTER GEOMETRY SECTION (" + where(c) + ")");
  }
}

}  // namespace necgpu
