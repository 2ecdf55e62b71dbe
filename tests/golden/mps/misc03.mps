*NAME:         misc03
*ROWS:         96
*COLUMNS:      160
*INTEGER:      159
*NONZERO:      2053
*BEST SOLN:    3360 (opt)
*LP SOLN:      1910.0
*SOURCE:   
*      
*              Greg Astfalk (Convex)
*APPLICATION:  unknown
*COMMENTS:     all integer variables are binary
*      
*      
NAME          MISC03
ROWS
 N  ROW097  
 E  ROW001  
 E  ROW002  
 E  ROW003  
 E  ROW004  
 E  ROW005  
 E  ROW006  
 L  ROW007  
 L  ROW008  
 L  ROW009  
 G  ROW010  
 G  ROW011  
 G  ROW012  
 G  ROW013  
 G  ROW014  
 G  ROW015  
 G  ROW016  
 G  ROW017  
 G  ROW018  
 G  ROW019  
 G  ROW020  
 G  ROW021  
 G  ROW022  
 G  ROW023  
 G  ROW024  
 G  ROW025  
 G  ROW026  
 G  ROW027  
 G  ROW028  
 G  ROW029  
 G  ROW030  
 G  ROW031  
 G  ROW032  
 G  ROW033  
 G  ROW034  
 G  ROW035  
 G  ROW036  
 G  ROW037  
 G  ROW038  
 G  ROW039  
 G  ROW040  
 E  ROW041  
 E  ROW042  
 E  ROW043  
 E  ROW044  
 E  ROW045  
 E  ROW046  
 E  ROW047  
 E  ROW048  
 E  ROW049  
 E  ROW050  
 E  ROW051  
 E  ROW052  
 E  ROW053  
 E  ROW054  
 E  ROW055  
 E  ROW056  
 E  ROW057  
 E  ROW058  
 E  ROW059  
 E  ROW060  
 E  ROW061  
 L  ROW062  
 L  ROW063  
 L  ROW064  
 L  ROW065  
 L  ROW066  
 L  ROW067  
 L  ROW068  
 L  ROW069  
 L  ROW070  
 L  ROW071  
 L  ROW072  
 L  ROW073  
 L  ROW074  
 L  ROW075  
 L  ROW076  
 L  ROW077  
 L  ROW078  
 L  ROW079  
 L  ROW080  
 L  ROW081  
 L  ROW082  
 L  ROW083  
 L  ROW084  
 L  ROW085  
 L  ROW086  
 L  ROW087  
 L  ROW088  
 L  ROW089  
 L  ROW090  
 L  ROW091  
 L  ROW092  
 L  ROW093  
 L  ROW094  
 L  ROW095  
 L  ROW096  
COLUMNS
    MARK0000  'MARKER'                 'INTORG'
    COL001    ROW002               1   ROW007             100
    COL001    ROW067               1   ROW070               1
    COL002    ROW002               1   ROW008             100
    COL002    ROW068               1   ROW071               1
    COL003    ROW002               1   ROW009             100
    COL003    ROW069               1   ROW072               1
    COL004    ROW001            -205   ROW002               1
    COL004    ROW007             100   ROW010               1
    COL004    ROW014               1   ROW018               1
    COL004    ROW022               1   ROW026               1
    COL004    ROW030               1   ROW034               1
    COL004    ROW038               1   ROW041               1
    COL004    ROW042              -1   ROW067               1
    COL004    ROW070               1
    COL005    ROW001            -205   ROW002               1
    COL005    ROW008             100   ROW010               1
    COL005    ROW014               1   ROW018               1
    COL005    ROW022               1   ROW026               1
    COL005    ROW030               1   ROW034               1
    COL005    ROW038               1   ROW048               1
    COL005    ROW049              -1   ROW068               1
    COL005    ROW071               1
    COL006    ROW001            -205   ROW002               1
    COL006    ROW009             100   ROW010               1
    COL006    ROW014               1   ROW018               1
    COL006    ROW022               1   ROW026               1
    COL006    ROW030               1   ROW034               1
    COL006    ROW038               1   ROW055               1
    COL006    ROW056              -1   ROW069               1
    COL006    ROW072               1
    COL007    ROW001            -550   ROW002               1
    COL007    ROW007             100   ROW010               1
    COL007    ROW012               1   ROW018               1
    COL007    ROW020               1   ROW026               1
    COL007    ROW028               1   ROW034               1
    COL007    ROW036               1   ROW041               1
    COL007    ROW043              -1   ROW067               1
    COL007    ROW070               1
    COL008    ROW001            -550   ROW002               1
    COL008    ROW008             100   ROW010               1
    COL008    ROW012               1   ROW018               1
    COL008    ROW020               1   ROW026               1
    COL008    ROW028               1   ROW034               1
    COL008    ROW036               1   ROW048               1
    COL008    ROW050              -1   ROW068               1
    COL008    ROW071               1
    COL009    ROW001            -550   ROW002               1
    COL009    ROW009             100   ROW010               1
    COL009    ROW012               1   ROW018               1
    COL009    ROW020               1   ROW026               1
    COL009    ROW028               1   ROW034               1
    COL009    ROW036               1   ROW055               1
    COL009    ROW057              -1   ROW069               1
    COL009    ROW072               1
    COL010    ROW001            -490   ROW002               1
    COL010    ROW007             100   ROW010               1
    COL010    ROW012               1   ROW014               1
    COL010    ROW016               1   ROW026               1
    COL010    ROW028               1   ROW030               1
    COL010    ROW032               1   ROW041               1
    COL010    ROW044              -1   ROW067               1
    COL010    ROW070               1
    COL011    ROW001            -490   ROW002               1
    COL011    ROW008             100   ROW010               1
    COL011    ROW012               1   ROW014               1
    COL011    ROW016               1   ROW026               1
    COL011    ROW028               1   ROW030               1
    COL011    ROW032               1   ROW048               1
    COL011    ROW051              -1   ROW068               1
    COL011    ROW071               1
    COL012    ROW001            -490   ROW002               1
    COL012    ROW009             100   ROW010               1
    COL012    ROW012               1   ROW014               1
    COL012    ROW016               1   ROW026               1
    COL012    ROW028               1   ROW030               1
    COL012    ROW032               1   ROW055               1
    COL012    ROW058              -1   ROW069               1
    COL012    ROW072               1
    COL013    ROW001            -660   ROW002               1
    COL013    ROW007             100   ROW010               1
    COL013    ROW012               1   ROW014               1
    COL013    ROW016               1   ROW018               1
    COL013    ROW020               1   ROW022               1
    COL013    ROW024               1   ROW041               1
    COL013    ROW045              -1   ROW067               1
    COL013    ROW070               1
    COL014    ROW001            -660   ROW002               1
    COL014    ROW008             100   ROW010               1
    COL014    ROW012               1   ROW014               1
    COL014    ROW016               1   ROW018               1
    COL014    ROW020               1   ROW022               1
    COL014    ROW024               1   ROW048               1
    COL014    ROW052              -1   ROW068               1
    COL014    ROW071               1
    COL015    ROW001            -660   ROW002               1
    COL015    ROW009             100   ROW010               1
    COL015    ROW012               1   ROW014               1
    COL015    ROW016               1   ROW018               1
    COL015    ROW020               1   ROW022               1
    COL015    ROW024               1   ROW055               1
    COL015    ROW059              -1   ROW069               1
    COL015    ROW072               1
    COL016    ROW001            -425   ROW002               1
    COL016    ROW007             100   ROW041               1
    COL016    ROW046              -1   ROW062               1
    COL016    ROW067               1   ROW070               1
    COL017    ROW001            -425   ROW002               1
    COL017    ROW008             100   ROW048               1
    COL017    ROW053              -1   ROW063               1
    COL017    ROW068               1   ROW071               1
    COL018    ROW001            -425   ROW002               1
    COL018    ROW009             100   ROW055               1
    COL018    ROW060              -1   ROW064               1
    COL018    ROW069               1   ROW072               1
    COL019    ROW001            -425   ROW002               1
    COL019    ROW007             100   ROW041               1
    COL019    ROW047              -1   ROW062               1
    COL019    ROW067               1   ROW070               1
    COL020    ROW001            -425   ROW002               1
    COL020    ROW008             100   ROW048               1
    COL020    ROW054              -1   ROW063               1
    COL020    ROW068               1   ROW071               1
    COL021    ROW001            -425   ROW002               1
    COL021    ROW009             100   ROW055               1
    COL021    ROW061              -1   ROW064               1
    COL021    ROW069               1   ROW072               1
    COL022    ROW001            -205   ROW003               1
    COL022    ROW007             100   ROW011               1
    COL022    ROW015               1   ROW019               1
    COL022    ROW023               1   ROW027               1
    COL022    ROW031               1   ROW035               1
    COL022    ROW039               1   ROW041              -1
    COL022    ROW042               1   ROW073               1
    COL022    ROW076               1
    COL023    ROW001            -205   ROW003               1
    COL023    ROW008             100   ROW011               1
    COL023    ROW015               1   ROW019               1
    COL023    ROW023               1   ROW027               1
    COL023    ROW031               1   ROW035               1
    COL023    ROW039               1   ROW048              -1
    COL023    ROW049               1   ROW074               1
    COL023    ROW077               1
    COL024    ROW001            -205   ROW003               1
    COL024    ROW009             100   ROW011               1
    COL024    ROW015               1   ROW019               1
    COL024    ROW023               1   ROW027               1
    COL024    ROW031               1   ROW035               1
    COL024    ROW039               1   ROW055              -1
    COL024    ROW056               1   ROW075               1
    COL024    ROW078               1
    COL025    ROW003               1   ROW007             100
    COL025    ROW073               1   ROW076               1
    COL026    ROW003               1   ROW008             100
    COL026    ROW074               1   ROW077               1
    COL027    ROW003               1   ROW009             100
    COL027    ROW075               1   ROW078               1
    COL028    ROW001            -350   ROW003               1
    COL028    ROW007             100   ROW011               1
    COL028    ROW012               1   ROW019               1
    COL028    ROW020               1   ROW027               1
    COL028    ROW028               1   ROW035               1
    COL028    ROW036               1   ROW042               1
    COL028    ROW043              -1   ROW073               1
    COL028    ROW076               1
    COL029    ROW001            -350   ROW003               1
    COL029    ROW008             100   ROW011               1
    COL029    ROW012               1   ROW019               1
    COL029    ROW020               1   ROW027               1
    COL029    ROW028               1   ROW035               1
    COL029    ROW036               1   ROW049               1
    COL029    ROW050              -1   ROW074               1
    COL029    ROW077               1
    COL030    ROW001            -350   ROW003               1
    COL030    ROW009             100   ROW011               1
    COL030    ROW012               1   ROW019               1
    COL030    ROW020               1   ROW027               1
    COL030    ROW028               1   ROW035               1
    COL030    ROW036               1   ROW056               1
    COL030    ROW057              -1   ROW075               1
    COL030    ROW078               1
    COL031    ROW001            -500   ROW003               1
    COL031    ROW007             100   ROW011               1
    COL031    ROW012               1   ROW015               1
    COL031    ROW016               1   ROW027               1
    COL031    ROW028               1   ROW031               1
    COL031    ROW032               1   ROW042               1
    COL031    ROW044              -1   ROW073               1
    COL031    ROW076               1
    COL032    ROW001            -500   ROW003               1
    COL032    ROW008             100   ROW011               1
    COL032    ROW012               1   ROW015               1
    COL032    ROW016               1   ROW027               1
    COL032    ROW028               1   ROW031               1
    COL032    ROW032               1   ROW049               1
    COL032    ROW051              -1   ROW074               1
    COL032    ROW077               1
    COL033    ROW001            -500   ROW003               1
    COL033    ROW009             100   ROW011               1
    COL033    ROW012               1   ROW015               1
    COL033    ROW016               1   ROW027               1
    COL033    ROW028               1   ROW031               1
    COL033    ROW032               1   ROW056               1
    COL033    ROW058              -1   ROW075               1
    COL033    ROW078               1
    COL034    ROW001            -760   ROW003               1
    COL034    ROW007             100   ROW011               1
    COL034    ROW012               1   ROW015               1
    COL034    ROW016               1   ROW019               1
    COL034    ROW020               1   ROW023               1
    COL034    ROW024               1   ROW042               1
    COL034    ROW045              -1   ROW073               1
    COL034    ROW076               1
    COL035    ROW001            -760   ROW003               1
    COL035    ROW008             100   ROW011               1
    COL035    ROW012               1   ROW015               1
    COL035    ROW016               1   ROW019               1
    COL035    ROW020               1   ROW023               1
    COL035    ROW024               1   ROW049               1
    COL035    ROW052              -1   ROW074               1
    COL035    ROW077               1
    COL036    ROW001            -760   ROW003               1
    COL036    ROW009             100   ROW011               1
    COL036    ROW012               1   ROW015               1
    COL036    ROW016               1   ROW019               1
    COL036    ROW020               1   ROW023               1
    COL036    ROW024               1   ROW056               1
    COL036    ROW059              -1   ROW075               1
    COL036    ROW078               1
    COL037    ROW001            -570   ROW003               1
    COL037    ROW007             100   ROW042               1
    COL037    ROW046              -1   ROW062               1
    COL037    ROW073               1   ROW076               1
    COL038    ROW001            -570   ROW003               1
    COL038    ROW008             100   ROW049               1
    COL038    ROW053              -1   ROW063               1
    COL038    ROW074               1   ROW077               1
    COL039    ROW001            -570   ROW003               1
    COL039    ROW009             100   ROW056               1
    COL039    ROW060              -1   ROW064               1
    COL039    ROW075               1   ROW078               1
    COL040    ROW001            -285   ROW003               1
    COL040    ROW007             100   ROW042               1
    COL040    ROW047              -1   ROW062               1
    COL040    ROW073               1   ROW076               1
    COL041    ROW001            -285   ROW003               1
    COL041    ROW008             100   ROW049               1
    COL041    ROW054              -1   ROW063               1
    COL041    ROW074               1   ROW077               1
    COL042    ROW001            -285   ROW003               1
    COL042    ROW009             100   ROW056               1
    COL042    ROW061              -1   ROW064               1
    COL042    ROW075               1   ROW078               1
    COL043    ROW001            -550   ROW004               1
    COL043    ROW007             100   ROW013               1
    COL043    ROW015               1   ROW021               1
    COL043    ROW023               1   ROW029               1
    COL043    ROW031               1   ROW037               1
    COL043    ROW039               1   ROW041              -1
    COL043    ROW043               1   ROW079               1
    COL043    ROW082               1
    COL044    ROW001            -550   ROW004               1
    COL044    ROW008             100   ROW013               1
    COL044    ROW015               1   ROW021               1
    COL044    ROW023               1   ROW029               1
    COL044    ROW031               1   ROW037               1
    COL044    ROW039               1   ROW048              -1
    COL044    ROW050               1   ROW080               1
    COL044    ROW083               1
    COL045    ROW001            -550   ROW004               1
    COL045    ROW009             100   ROW013               1
    COL045    ROW015               1   ROW021               1
    COL045    ROW023               1   ROW029               1
    COL045    ROW031               1   ROW037               1
    COL045    ROW039               1   ROW055              -1
    COL045    ROW057               1   ROW081               1
    COL045    ROW084               1
    COL046    ROW001            -350   ROW004               1
    COL046    ROW007             100   ROW013               1
    COL046    ROW014               1   ROW021               1
    COL046    ROW022               1   ROW029               1
    COL046    ROW030               1   ROW037               1
    COL046    ROW038               1   ROW042              -1
    COL046    ROW043               1   ROW079               1
    COL046    ROW082               1
    COL047    ROW001            -350   ROW004               1
    COL047    ROW008             100   ROW013               1
    COL047    ROW014               1   ROW021               1
    COL047    ROW022               1   ROW029               1
    COL047    ROW030               1   ROW037               1
    COL047    ROW038               1   ROW049              -1
    COL047    ROW050               1   ROW080               1
    COL047    ROW083               1
    COL048    ROW001            -350   ROW004               1
    COL048    ROW009             100   ROW013               1
    COL048    ROW014               1   ROW021               1
    COL048    ROW022               1   ROW029               1
    COL048    ROW030               1   ROW037               1
    COL048    ROW038               1   ROW056              -1
    COL048    ROW057               1   ROW081               1
    COL048    ROW084               1
    COL049    ROW004               1   ROW007             100
    COL049    ROW079               1   ROW082               1
    COL050    ROW004               1   ROW008             100
    COL050    ROW080               1   ROW083               1
    COL051    ROW004               1   ROW009             100
    COL051    ROW081               1   ROW084               1
    COL052    ROW001            -610   ROW004               1
    COL052    ROW007             100   ROW013               1
    COL052    ROW014               1   ROW015               1
    COL052    ROW016               1   ROW029               1
    COL052    ROW030               1   ROW031               1
    COL052    ROW032               1   ROW043               1
    COL052    ROW044              -1   ROW079               1
    COL052    ROW082               1
    COL053    ROW001            -610   ROW004               1
    COL053    ROW008             100   ROW013               1
    COL053    ROW014               1   ROW015               1
    COL053    ROW016               1   ROW029               1
    COL053    ROW030               1   ROW031               1
    COL053    ROW032               1   ROW050               1
    COL053    ROW051              -1   ROW080               1
    COL053    ROW083               1
    COL054    ROW001            -610   ROW004               1
    COL054    ROW009             100   ROW013               1
    COL054    ROW014               1   ROW015               1
    COL054    ROW016               1   ROW029               1
    COL054    ROW030               1   ROW031               1
    COL054    ROW032               1   ROW057               1
    COL054    ROW058              -1   ROW081               1
    COL054    ROW084               1
    COL055    ROW001            -955   ROW004               1
    COL055    ROW007             100   ROW013               1
    COL055    ROW014               1   ROW015               1
    COL055    ROW016               1   ROW021               1
    COL055    ROW022               1   ROW023               1
    COL055    ROW024               1   ROW043               1
    COL055    ROW045              -1   ROW079               1
    COL055    ROW082               1
    COL056    ROW001            -955   ROW004               1
    COL056    ROW008             100   ROW013               1
    COL056    ROW014               1   ROW015               1
    COL056    ROW016               1   ROW021               1
    COL056    ROW022               1   ROW023               1
    COL056    ROW024               1   ROW050               1
    COL056    ROW052              -1   ROW080               1
    COL056    ROW083               1
    COL057    ROW001            -955   ROW004               1
    COL057    ROW009             100   ROW013               1
    COL057    ROW014               1   ROW015               1
    COL057    ROW016               1   ROW021               1
    COL057    ROW022               1   ROW023               1
    COL057    ROW024               1   ROW057               1
    COL057    ROW059              -1   ROW081               1
    COL057    ROW084               1
    COL058    ROW001            -830   ROW004               1
    COL058    ROW007             100   ROW043               1
    COL058    ROW046              -1   ROW062               1
    COL058    ROW079               1   ROW082               1
    COL059    ROW001            -830   ROW004               1
    COL059    ROW008             100   ROW050               1
    COL059    ROW053              -1   ROW063               1
    COL059    ROW080               1   ROW083               1
    COL060    ROW001            -830   ROW004               1
    COL060    ROW009             100   ROW057               1
    COL060    ROW060              -1   ROW064               1
    COL060    ROW081               1   ROW084               1
    COL061    ROW001            -250   ROW004               1
    COL061    ROW007             100   ROW043               1
    COL061    ROW047              -1   ROW062               1
    COL061    ROW079               1   ROW082               1
    COL062    ROW001            -250   ROW004               1
    COL062    ROW008             100   ROW050               1
    COL062    ROW054              -1   ROW063               1
    COL062    ROW080               1   ROW083               1
    COL063    ROW001            -250   ROW004               1
    COL063    ROW009             100   ROW057               1
    COL063    ROW061              -1   ROW064               1
    COL063    ROW081               1   ROW084               1
    COL064    ROW001            -490   ROW005               1
    COL064    ROW007             100   ROW017               1
    COL064    ROW019               1   ROW021               1
    COL064    ROW023               1   ROW033               1
    COL064    ROW035               1   ROW037               1
    COL064    ROW039               1   ROW041              -1
    COL064    ROW044               1   ROW085               1
    COL064    ROW088               1
    COL065    ROW001            -490   ROW005               1
    COL065    ROW008             100   ROW017               1
    COL065    ROW019               1   ROW021               1
    COL065    ROW023               1   ROW033               1
    COL065    ROW035               1   ROW037               1
    COL065    ROW039               1   ROW048              -1
    COL065    ROW051               1   ROW086               1
    COL065    ROW089               1
    COL066    ROW001            -490   ROW005               1
    COL066    ROW009             100   ROW017               1
    COL066    ROW019               1   ROW021               1
    COL066    ROW023               1   ROW033               1
    COL066    ROW035               1   ROW037               1
    COL066    ROW039               1   ROW055              -1
    COL066    ROW058               1   ROW087               1
    COL066    ROW090               1
    COL067    ROW001            -500   ROW005               1
    COL067    ROW007             100   ROW017               1
    COL067    ROW018               1   ROW021               1
    COL067    ROW022               1   ROW033               1
    COL067    ROW034               1   ROW037               1
    COL067    ROW038               1   ROW042              -1
    COL067    ROW044               1   ROW085               1
    COL067    ROW088               1
    COL068    ROW001            -500   ROW005               1
    COL068    ROW008             100   ROW017               1
    COL068    ROW018               1   ROW021               1
    COL068    ROW022               1   ROW033               1
    COL068    ROW034               1   ROW037               1
    COL068    ROW038               1   ROW049              -1
    COL068    ROW051               1   ROW086               1
    COL068    ROW089               1
    COL069    ROW001            -500   ROW005               1
    COL069    ROW009             100   ROW017               1
    COL069    ROW018               1   ROW021               1
    COL069    ROW022               1   ROW033               1
    COL069    ROW034               1   ROW037               1
    COL069    ROW038               1   ROW056              -1
    COL069    ROW058               1   ROW087               1
    COL069    ROW090               1
    COL070    ROW001            -610   ROW005               1
    COL070    ROW007             100   ROW017               1
    COL070    ROW018               1   ROW019               1
    COL070    ROW020               1   ROW033               1
    COL070    ROW034               1   ROW035               1
    COL070    ROW036               1   ROW043              -1
    COL070    ROW044               1   ROW085               1
    COL070    ROW088               1
    COL071    ROW001            -610   ROW005               1
    COL071    ROW008             100   ROW017               1
    COL071    ROW018               1   ROW019               1
    COL071    ROW020               1   ROW033               1
    COL071    ROW034               1   ROW035               1
    COL071    ROW036               1   ROW050              -1
    COL071    ROW051               1   ROW086               1
    COL071    ROW089               1
    COL072    ROW001            -610   ROW005               1
    COL072    ROW009             100   ROW017               1
    COL072    ROW018               1   ROW019               1
    COL072    ROW020               1   ROW033               1
    COL072    ROW034               1   ROW035               1
    COL072    ROW036               1   ROW057              -1
    COL072    ROW058               1   ROW087               1
    COL072    ROW090               1
    COL073    ROW005               1   ROW007             100
    COL073    ROW085               1   ROW088               1
    COL074    ROW005               1   ROW008             100
    COL074    ROW086               1   ROW089               1
    COL075    ROW005               1   ROW009             100
    COL075    ROW087               1   ROW090               1
    COL076    ROW001            -360   ROW005               1
    COL076    ROW007             100   ROW017               1
    COL076    ROW018               1   ROW019               1
    COL076    ROW020               1   ROW021               1
    COL076    ROW022               1   ROW023               1
    COL076    ROW024               1   ROW044               1
    COL076    ROW045              -1   ROW085               1
    COL076    ROW088               1
    COL077    ROW001            -360   ROW005               1
    COL077    ROW008             100   ROW017               1
    COL077    ROW018               1   ROW019               1
    COL077    ROW020               1   ROW021               1
    COL077    ROW022               1   ROW023               1
    COL077    ROW024               1   ROW051               1
    COL077    ROW052              -1   ROW086               1
    COL077    ROW089               1
    COL078    ROW001            -360   ROW005               1
    COL078    ROW009             100   ROW017               1
    COL078    ROW018               1   ROW019               1
    COL078    ROW020               1   ROW021               1
    COL078    ROW022               1   ROW023               1
    COL078    ROW024               1   ROW058               1
    COL078    ROW059              -1   ROW087               1
    COL078    ROW090               1
    COL079    ROW001            -355   ROW005               1
    COL079    ROW007             100   ROW044               1
    COL079    ROW046              -1   ROW062               1
    COL079    ROW085               1   ROW088               1
    COL080    ROW001            -355   ROW005               1
    COL080    ROW008             100   ROW051               1
    COL080    ROW053              -1   ROW063               1
    COL080    ROW086               1   ROW089               1
    COL081    ROW001            -355   ROW005               1
    COL081    ROW009             100   ROW058               1
    COL081    ROW060              -1   ROW064               1
    COL081    ROW087               1   ROW090               1
    COL082    ROW001            -360   ROW005               1
    COL082    ROW007             100   ROW044               1
    COL082    ROW047              -1   ROW062               1
    COL082    ROW085               1   ROW088               1
    COL083    ROW001            -360   ROW005               1
    COL083    ROW008             100   ROW051               1
    COL083    ROW054              -1   ROW063               1
    COL083    ROW086               1   ROW089               1
    COL084    ROW001            -360   ROW005               1
    COL084    ROW009             100   ROW058               1
    COL084    ROW061              -1   ROW064               1
    COL084    ROW087               1   ROW090               1
    COL085    ROW001            -660   ROW006               1
    COL085    ROW007             100   ROW025               1
    COL085    ROW027               1   ROW029               1
    COL085    ROW031               1   ROW033               1
    COL085    ROW035               1   ROW037               1
    COL085    ROW039               1   ROW041              -1
    COL085    ROW045               1   ROW091               1
    COL085    ROW094               1
    COL086    ROW001            -660   ROW006               1
    COL086    ROW008             100   ROW025               1
    COL086    ROW027               1   ROW029               1
    COL086    ROW031               1   ROW033               1
    COL086    ROW035               1   ROW037               1
    COL086    ROW039               1   ROW048              -1
    COL086    ROW052               1   ROW092               1
    COL086    ROW095               1
    COL087    ROW001            -660   ROW006               1
    COL087    ROW009             100   ROW025               1
    COL087    ROW027               1   ROW029               1
    COL087    ROW031               1   ROW033               1
    COL087    ROW035               1   ROW037               1
    COL087    ROW039               1   ROW055              -1
    COL087    ROW059               1   ROW093               1
    COL087    ROW096               1
    COL088    ROW001            -760   ROW006               1
    COL088    ROW007             100   ROW025               1
    COL088    ROW026               1   ROW029               1
    COL088    ROW030               1   ROW033               1
    COL088    ROW034               1   ROW037               1
    COL088    ROW038               1   ROW042              -1
    COL088    ROW045               1   ROW091               1
    COL088    ROW094               1
    COL089    ROW001            -760   ROW006               1
    COL089    ROW008             100   ROW025               1
    COL089    ROW026               1   ROW029               1
    COL089    ROW030               1   ROW033               1
    COL089    ROW034               1   ROW037               1
    COL089    ROW038               1   ROW049              -1
    COL089    ROW052               1   ROW092               1
    COL089    ROW095               1
    COL090    ROW001            -760   ROW006               1
    COL090    ROW009             100   ROW025               1
    COL090    ROW026               1   ROW029               1
    COL090    ROW030               1   ROW033               1
    COL090    ROW034               1   ROW037               1
    COL090    ROW038               1   ROW056              -1
    COL090    ROW059               1   ROW093               1
    COL090    ROW096               1
    COL091    ROW001            -955   ROW006               1
    COL091    ROW007             100   ROW025               1
    COL091    ROW026               1   ROW027               1
    COL091    ROW028               1   ROW033               1
    COL091    ROW034               1   ROW035               1
    COL091    ROW036               1   ROW043              -1
    COL091    ROW045               1   ROW091               1
    COL091    ROW094               1
    COL092    ROW001            -955   ROW006               1
    COL092    ROW008             100   ROW025               1
    COL092    ROW026               1   ROW027               1
    COL092    ROW028               1   ROW033               1
    COL092    ROW034               1   ROW035               1
    COL092    ROW036               1   ROW050              -1
    COL092    ROW052               1   ROW092               1
    COL092    ROW095               1
    COL093    ROW001            -955   ROW006               1
    COL093    ROW009             100   ROW025               1
    COL093    ROW026               1   ROW027               1
    COL093    ROW028               1   ROW033               1
    COL093    ROW034               1   ROW035               1
    COL093    ROW036               1   ROW057              -1
    COL093    ROW059               1   ROW093               1
    COL093    ROW096               1
    COL094    ROW001            -360   ROW006               1
    COL094    ROW007             100   ROW025               1
    COL094    ROW026               1   ROW027               1
    COL094    ROW028               1   ROW029               1
    COL094    ROW030               1   ROW031               1
    COL094    ROW032               1   ROW044              -1
    COL094    ROW045               1   ROW091               1
    COL094    ROW094               1
    COL095    ROW001            -360   ROW006               1
    COL095    ROW008             100   ROW025               1
    COL095    ROW026               1   ROW027               1
    COL095    ROW028               1   ROW029               1
    COL095    ROW030               1   ROW031               1
    COL095    ROW032               1   ROW051              -1
    COL095    ROW052               1   ROW092               1
    COL095    ROW095               1
    COL096    ROW001            -360   ROW006               1
    COL096    ROW009             100   ROW025               1
    COL096    ROW026               1   ROW027               1
    COL096    ROW028               1   ROW029               1
    COL096    ROW030               1   ROW031               1
    COL096    ROW032               1   ROW058              -1
    COL096    ROW059               1   ROW093               1
    COL096    ROW096               1
    COL097    ROW006               1   ROW007             100
    COL097    ROW091               1   ROW094               1
    COL098    ROW006               1   ROW008             100
    COL098    ROW092               1   ROW095               1
    COL099    ROW006               1   ROW009             100
    COL099    ROW093               1   ROW096               1
    COL100    ROW001            -255   ROW006               1
    COL100    ROW007             100   ROW045               1
    COL100    ROW046              -1   ROW062               1
    COL100    ROW091               1   ROW094               1
    COL101    ROW001            -255   ROW006               1
    COL101    ROW008             100   ROW052               1
    COL101    ROW053              -1   ROW063               1
    COL101    ROW092               1   ROW095               1
    COL102    ROW001            -255   ROW006               1
    COL102    ROW009             100   ROW059               1
    COL102    ROW060              -1   ROW064               1
    COL102    ROW093               1   ROW096               1
    COL103    ROW001            -705   ROW006               1
    COL103    ROW007             100   ROW045               1
    COL103    ROW047              -1   ROW062               1
    COL103    ROW091               1   ROW094               1
    COL104    ROW001            -705   ROW006               1
    COL104    ROW008             100   ROW052               1
    COL104    ROW054              -1   ROW063               1
    COL104    ROW092               1   ROW095               1
    COL105    ROW001            -705   ROW006               1
    COL105    ROW009             100   ROW059               1
    COL105    ROW061              -1   ROW064               1
    COL105    ROW093               1   ROW096               1
    COL106    ROW001            -425   ROW011               1
    COL106    ROW013               1   ROW015               1
    COL106    ROW017               1   ROW019               1
    COL106    ROW021               1   ROW023               1
    COL106    ROW025               1   ROW027               1
    COL106    ROW029               1   ROW031               1
    COL106    ROW033               1   ROW035               1
    COL106    ROW037               1   ROW039               1
    COL106    ROW040               1   ROW041              -1
    COL106    ROW046               1   ROW067               1
    COL106    ROW073               1   ROW079               1
    COL106    ROW085               1   ROW091               1
    COL107    ROW001            -425   ROW011               1
    COL107    ROW013               1   ROW015               1
    COL107    ROW017               1   ROW019               1
    COL107    ROW021               1   ROW023               1
    COL107    ROW025               1   ROW027               1
    COL107    ROW029               1   ROW031               1
    COL107    ROW033               1   ROW035               1
    COL107    ROW037               1   ROW039               1
    COL107    ROW040               1   ROW048              -1
    COL107    ROW053               1   ROW068               1
    COL107    ROW074               1   ROW080               1
    COL107    ROW086               1   ROW092               1
    COL108    ROW001            -425   ROW011               1
    COL108    ROW013               1   ROW015               1
    COL108    ROW017               1   ROW019               1
    COL108    ROW021               1   ROW023               1
    COL108    ROW025               1   ROW027               1
    COL108    ROW029               1   ROW031               1
    COL108    ROW033               1   ROW035               1
    COL108    ROW037               1   ROW039               1
    COL108    ROW040               1   ROW055              -1
    COL108    ROW060               1   ROW069               1
    COL108    ROW075               1   ROW081               1
    COL108    ROW087               1   ROW093               1
    COL109    ROW001            -570   ROW010               1
    COL109    ROW013               1   ROW014               1
    COL109    ROW017               1   ROW018               1
    COL109    ROW021               1   ROW022               1
    COL109    ROW025               1   ROW026               1
    COL109    ROW029               1   ROW030               1
    COL109    ROW033               1   ROW034               1
    COL109    ROW037               1   ROW038               1
    COL109    ROW040               1   ROW042              -1
    COL109    ROW046               1   ROW067               1
    COL109    ROW073               1   ROW079               1
    COL109    ROW085               1   ROW091               1
    COL110    ROW001            -570   ROW010               1
    COL110    ROW013               1   ROW014               1
    COL110    ROW017               1   ROW018               1
    COL110    ROW021               1   ROW022               1
    COL110    ROW025               1   ROW026               1
    COL110    ROW029               1   ROW030               1
    COL110    ROW033               1   ROW034               1
    COL110    ROW037               1   ROW038               1
    COL110    ROW040               1   ROW049              -1
    COL110    ROW053               1   ROW068               1
    COL110    ROW074               1   ROW080               1
    COL110    ROW086               1   ROW092               1
    COL111    ROW001            -570   ROW010               1
    COL111    ROW013               1   ROW014               1
    COL111    ROW017               1   ROW018               1
    COL111    ROW021               1   ROW022               1
    COL111    ROW025               1   ROW026               1
    COL111    ROW029               1   ROW030               1
    COL111    ROW033               1   ROW034               1
    COL111    ROW037               1   ROW038               1
    COL111    ROW040               1   ROW056              -1
    COL111    ROW060               1   ROW069               1
    COL111    ROW075               1   ROW081               1
    COL111    ROW087               1   ROW093               1
    COL112    ROW001            -830   ROW010               1
    COL112    ROW011               1   ROW012               1
    COL112    ROW017               1   ROW018               1
    COL112    ROW019               1   ROW020               1
    COL112    ROW025               1   ROW026               1
    COL112    ROW027               1   ROW028               1
    COL112    ROW033               1   ROW034               1
    COL112    ROW035               1   ROW036               1
    COL112    ROW040               1   ROW043              -1
    COL112    ROW046               1   ROW067               1
    COL112    ROW073               1   ROW079               1
    COL112    ROW085               1   ROW091               1
    COL113    ROW001            -830   ROW010               1
    COL113    ROW011               1   ROW012               1
    COL113    ROW017               1   ROW018               1
    COL113    ROW019               1   ROW020               1
    COL113    ROW025               1   ROW026               1
    COL113    ROW027               1   ROW028               1
    COL113    ROW033               1   ROW034               1
    COL113    ROW035               1   ROW036               1
    COL113    ROW040               1   ROW050              -1
    COL113    ROW053               1   ROW068               1
    COL113    ROW074               1   ROW080               1
    COL113    ROW086               1   ROW092               1
    COL114    ROW001            -830   ROW010               1
    COL114    ROW011               1   ROW012               1
    COL114    ROW017               1   ROW018               1
    COL114    ROW019               1   ROW020               1
    COL114    ROW025               1   ROW026               1
    COL114    ROW027               1   ROW028               1
    COL114    ROW033               1   ROW034               1
    COL114    ROW035               1   ROW036               1
    COL114    ROW040               1   ROW057              -1
    COL114    ROW060               1   ROW069               1
    COL114    ROW075               1   ROW081               1
    COL114    ROW087               1   ROW093               1
    COL115    ROW001            -355   ROW010               1
    COL115    ROW011               1   ROW012               1
    COL115    ROW013               1   ROW014               1
    COL115    ROW015               1   ROW016               1
    COL115    ROW025               1   ROW026               1
    COL115    ROW027               1   ROW028               1
    COL115    ROW029               1   ROW030               1
    COL115    ROW031               1   ROW032               1
    COL115    ROW040               1   ROW044              -1
    COL115    ROW046               1   ROW067               1
    COL115    ROW073               1   ROW079               1
    COL115    ROW085               1   ROW091               1
    COL116    ROW001            -355   ROW010               1
    COL116    ROW011               1   ROW012               1
    COL116    ROW013               1   ROW014               1
    COL116    ROW015               1   ROW016               1
    COL116    ROW025               1   ROW026               1
    COL116    ROW027               1   ROW028               1
    COL116    ROW029               1   ROW030               1
    COL116    ROW031               1   ROW032               1
    COL116    ROW040               1   ROW051              -1
    COL116    ROW053               1   ROW068               1
    COL116    ROW074               1   ROW080               1
    COL116    ROW086               1   ROW092               1
    COL117    ROW001            -355   ROW010               1
    COL117    ROW011               1   ROW012               1
    COL117    ROW013               1   ROW014               1
    COL117    ROW015               1   ROW016               1
    COL117    ROW025               1   ROW026               1
    COL117    ROW027               1   ROW028               1
    COL117    ROW029               1   ROW030               1
    COL117    ROW031               1   ROW032               1
    COL117    ROW040               1   ROW058              -1
    COL117    ROW060               1   ROW069               1
    COL117    ROW075               1   ROW081               1
    COL117    ROW087               1   ROW093               1
    COL118    ROW001            -255   ROW010               1
    COL118    ROW011               1   ROW012               1
    COL118    ROW013               1   ROW014               1
    COL118    ROW015               1   ROW016               1
    COL118    ROW017               1   ROW018               1
    COL118    ROW019               1   ROW020               1
    COL118    ROW021               1   ROW022               1
    COL118    ROW023               1   ROW024               1
    COL118    ROW040               1   ROW045              -1
    COL118    ROW046               1   ROW067               1
    COL118    ROW073               1   ROW079               1
    COL118    ROW085               1   ROW091               1
    COL119    ROW001            -255   ROW010               1
    COL119    ROW011               1   ROW012               1
    COL119    ROW013               1   ROW014               1
    COL119    ROW015               1   ROW016               1
    COL119    ROW017               1   ROW018               1
    COL119    ROW019               1   ROW020               1
    COL119    ROW021               1   ROW022               1
    COL119    ROW023               1   ROW024               1
    COL119    ROW040               1   ROW052              -1
    COL119    ROW053               1   ROW068               1
    COL119    ROW074               1   ROW080               1
    COL119    ROW086               1   ROW092               1
    COL120    ROW001            -255   ROW010               1
    COL120    ROW011               1   ROW012               1
    COL120    ROW013               1   ROW014               1
    COL120    ROW015               1   ROW016               1
    COL120    ROW017               1   ROW018               1
    COL120    ROW019               1   ROW020               1
    COL120    ROW021               1   ROW022               1
    COL120    ROW023               1   ROW024               1
    COL120    ROW040               1   ROW059              -1
    COL120    ROW060               1   ROW069               1
    COL120    ROW075               1   ROW081               1
    COL120    ROW087               1   ROW093               1
    COL121    ROW067               1   ROW073               1
    COL121    ROW079               1   ROW085               1
    COL121    ROW091               1
    COL122    ROW068               1   ROW074               1
    COL122    ROW080               1   ROW086               1
    COL122    ROW092               1
    COL123    ROW069               1   ROW075               1
    COL123    ROW081               1   ROW087               1
    COL123    ROW093               1
    COL124    ROW001            -605   ROW046               1
    COL124    ROW047              -1   ROW067               1
    COL124    ROW073               1   ROW079               1
    COL124    ROW085               1   ROW091               1
    COL125    ROW001            -605   ROW053               1
    COL125    ROW054              -1   ROW068               1
    COL125    ROW074               1   ROW080               1
    COL125    ROW086               1   ROW092               1
    COL126    ROW001            -605   ROW060               1
    COL126    ROW061              -1   ROW069               1
    COL126    ROW075               1   ROW081               1
    COL126    ROW087               1   ROW093               1
    COL127    ROW001            -425   ROW011               1
    COL127    ROW013               1   ROW015               1
    COL127    ROW017               1   ROW019               1
    COL127    ROW021               1   ROW023               1
    COL127    ROW025               1   ROW027               1
    COL127    ROW029               1   ROW031               1
    COL127    ROW033               1   ROW035               1
    COL127    ROW037               1   ROW039               1
    COL127    ROW040               1   ROW041              -1
    COL127    ROW047               1   ROW070               1
    COL127    ROW076               1   ROW082               1
    COL127    ROW088               1   ROW094               1
    COL128    ROW001            -425   ROW011               1
    COL128    ROW013               1   ROW015               1
    COL128    ROW017               1   ROW019               1
    COL128    ROW021               1   ROW023               1
    COL128    ROW025               1   ROW027               1
    COL128    ROW029               1   ROW031               1
    COL128    ROW033               1   ROW035               1
    COL128    ROW037               1   ROW039               1
    COL128    ROW040               1   ROW048              -1
    COL128    ROW054               1   ROW071               1
    COL128    ROW077               1   ROW083               1
    COL128    ROW089               1   ROW095               1
    COL129    ROW001            -425   ROW011               1
    COL129    ROW013               1   ROW015               1
    COL129    ROW017               1   ROW019               1
    COL129    ROW021               1   ROW023               1
    COL129    ROW025               1   ROW027               1
    COL129    ROW029               1   ROW031               1
    COL129    ROW033               1   ROW035               1
    COL129    ROW037               1   ROW039               1
    COL129    ROW040               1   ROW055              -1
    COL129    ROW061               1   ROW072               1
    COL129    ROW078               1   ROW084               1
    COL129    ROW090               1   ROW096               1
    COL130    ROW001            -285   ROW010               1
    COL130    ROW013               1   ROW014               1
    COL130    ROW017               1   ROW018               1
    COL130    ROW021               1   ROW022               1
    COL130    ROW025               1   ROW026               1
    COL130    ROW029               1   ROW030               1
    COL130    ROW033               1   ROW034               1
    COL130    ROW037               1   ROW038               1
    COL130    ROW040               1   ROW042              -1
    COL130    ROW047               1   ROW070               1
    COL130    ROW076               1   ROW082               1
    COL130    ROW088               1   ROW094               1
    COL131    ROW001            -285   ROW010               1
    COL131    ROW013               1   ROW014               1
    COL131    ROW017               1   ROW018               1
    COL131    ROW021               1   ROW022               1
    COL131    ROW025               1   ROW026               1
    COL131    ROW029               1   ROW030               1
    COL131    ROW033               1   ROW034               1
    COL131    ROW037               1   ROW038               1
    COL131    ROW040               1   ROW049              -1
    COL131    ROW054               1   ROW071               1
    COL131    ROW077               1   ROW083               1
    COL131    ROW089               1   ROW095               1
    COL132    ROW001            -285   ROW010               1
    COL132    ROW013               1   ROW014               1
    COL132    ROW017               1   ROW018               1
    COL132    ROW021               1   ROW022               1
    COL132    ROW025               1   ROW026               1
    COL132    ROW029               1   ROW030               1
    COL132    ROW033               1   ROW034               1
    COL132    ROW037               1   ROW038               1
    COL132    ROW040               1   ROW056              -1
    COL132    ROW061               1   ROW072               1
    COL132    ROW078               1   ROW084               1
    COL132    ROW090               1   ROW096               1
    COL133    ROW001            -250   ROW010               1
    COL133    ROW011               1   ROW012               1
    COL133    ROW017               1   ROW018               1
    COL133    ROW019               1   ROW020               1
    COL133    ROW025               1   ROW026               1
    COL133    ROW027               1   ROW028               1
    COL133    ROW033               1   ROW034               1
    COL133    ROW035               1   ROW036               1
    COL133    ROW040               1   ROW043              -1
    COL133    ROW047               1   ROW070               1
    COL133    ROW076               1   ROW082               1
    COL133    ROW088               1   ROW094               1
    COL134    ROW001            -250   ROW010               1
    COL134    ROW011               1   ROW012               1
    COL134    ROW017               1   ROW018               1
    COL134    ROW019               1   ROW020               1
    COL134    ROW025               1   ROW026               1
    COL134    ROW027               1   ROW028               1
    COL134    ROW033               1   ROW034               1
    COL134    ROW035               1   ROW036               1
    COL134    ROW040               1   ROW050              -1
    COL134    ROW054               1   ROW071               1
    COL134    ROW077               1   ROW083               1
    COL134    ROW089               1   ROW095               1
    COL135    ROW001            -250   ROW010               1
    COL135    ROW011               1   ROW012               1
    COL135    ROW017               1   ROW018               1
    COL135    ROW019               1   ROW020               1
    COL135    ROW025               1   ROW026               1
    COL135    ROW027               1   ROW028               1
    COL135    ROW033               1   ROW034               1
    COL135    ROW035               1   ROW036               1
    COL135    ROW040               1   ROW057              -1
    COL135    ROW061               1   ROW072               1
    COL135    ROW078               1   ROW084               1
    COL135    ROW090               1   ROW096               1
    COL136    ROW001            -360   ROW010               1
    COL136    ROW011               1   ROW012               1
    COL136    ROW013               1   ROW014               1
    COL136    ROW015               1   ROW016               1
    COL136    ROW025               1   ROW026               1
    COL136    ROW027               1   ROW028               1
    COL136    ROW029               1   ROW030               1
    COL136    ROW031               1   ROW032               1
    COL136    ROW040               1   ROW044              -1
    COL136    ROW047               1   ROW070               1
    COL136    ROW076               1   ROW082               1
    COL136    ROW088               1   ROW094               1
    COL137    ROW001            -360   ROW010               1
    COL137    ROW011               1   ROW012               1
    COL137    ROW013               1   ROW014               1
    COL137    ROW015               1   ROW016               1
    COL137    ROW025               1   ROW026               1
    COL137    ROW027               1   ROW028               1
    COL137    ROW029               1   ROW030               1
    COL137    ROW031               1   ROW032               1
    COL137    ROW040               1   ROW051              -1
    COL137    ROW054               1   ROW071               1
    COL137    ROW077               1   ROW083               1
    COL137    ROW089               1   ROW095               1
    COL138    ROW001            -360   ROW010               1
    COL138    ROW011               1   ROW012               1
    COL138    ROW013               1   ROW014               1
    COL138    ROW015               1   ROW016               1
    COL138    ROW025               1   ROW026               1
    COL138    ROW027               1   ROW028               1
    COL138    ROW029               1   ROW030               1
    COL138    ROW031               1   ROW032               1
    COL138    ROW040               1   ROW058              -1
    COL138    ROW061               1   ROW072               1
    COL138    ROW078               1   ROW084               1
    COL138    ROW090               1   ROW096               1
    COL139    ROW001            -705   ROW010               1
    COL139    ROW011               1   ROW012               1
    COL139    ROW013               1   ROW014               1
    COL139    ROW015               1   ROW016               1
    COL139    ROW017               1   ROW018               1
    COL139    ROW019               1   ROW020               1
    COL139    ROW021               1   ROW022               1
    COL139    ROW023               1   ROW024               1
    COL139    ROW040               1   ROW045              -1
    COL139    ROW047               1   ROW070               1
    COL139    ROW076               1   ROW082               1
    COL139    ROW088               1   ROW094               1
    COL140    ROW001            -705   ROW010               1
    COL140    ROW011               1   ROW012               1
    COL140    ROW013               1   ROW014               1
    COL140    ROW015               1   ROW016               1
    COL140    ROW017               1   ROW018               1
    COL140    ROW019               1   ROW020               1
    COL140    ROW021               1   ROW022               1
    COL140    ROW023               1   ROW024               1
    COL140    ROW040               1   ROW052              -1
    COL140    ROW054               1   ROW071               1
    COL140    ROW077               1   ROW083               1
    COL140    ROW089               1   ROW095               1
    COL141    ROW001            -705   ROW010               1
    COL141    ROW011               1   ROW012               1
    COL141    ROW013               1   ROW014               1
    COL141    ROW015               1   ROW016               1
    COL141    ROW017               1   ROW018               1
    COL141    ROW019               1   ROW020               1
    COL141    ROW021               1   ROW022               1
    COL141    ROW023               1   ROW024               1
    COL141    ROW040               1   ROW059              -1
    COL141    ROW061               1   ROW072               1
    COL141    ROW078               1   ROW084               1
    COL141    ROW090               1   ROW096               1
    COL142    ROW001            -605   ROW046              -1
    COL142    ROW047               1   ROW070               1
    COL142    ROW076               1   ROW082               1
    COL142    ROW088               1   ROW094               1
    COL143    ROW001            -605   ROW053              -1
    COL143    ROW054               1   ROW071               1
    COL143    ROW077               1   ROW083               1
    COL143    ROW089               1   ROW095               1
    COL144    ROW001            -605   ROW060              -1
    COL144    ROW061               1   ROW072               1
    COL144    ROW078               1   ROW084               1
    COL144    ROW090               1   ROW096               1
    COL145    ROW070               1   ROW076               1
    COL145    ROW082               1   ROW088               1
    COL145    ROW094               1
    COL146    ROW071               1   ROW077               1
    COL146    ROW083               1   ROW089               1
    COL146    ROW095               1
    COL147    ROW072               1   ROW078               1
    COL147    ROW084               1   ROW090               1
    COL147    ROW096               1
    COL148    ROW001            -300   ROW065            -300
    COL149    ROW001            -300   ROW066            -300
    COL150    ROW001             -75   ROW065             100
    COL150    ROW067              -1   ROW068              -1
    COL150    ROW069              -1
    COL151    ROW001             -75   ROW066             100
    COL151    ROW070              -1   ROW071              -1
    COL151    ROW072              -1
    COL152    ROW001             -75   ROW065             100
    COL152    ROW073              -1   ROW074              -1
    COL152    ROW075              -1
    COL153    ROW001             -75   ROW066             100
    COL153    ROW076              -1   ROW077              -1
    COL153    ROW078              -1
    COL154    ROW001             -75   ROW065             100
    COL154    ROW079              -1   ROW080              -1
    COL154    ROW081              -1
    COL155    ROW001             -75   ROW066             100
    COL155    ROW082              -1   ROW083              -1
    COL155    ROW084              -1
    COL156    ROW001             -75   ROW065             100
    COL156    ROW085              -1   ROW086              -1
    COL156    ROW087              -1
    COL157    ROW001             -75   ROW066             100
    COL157    ROW088              -1   ROW089              -1
    COL157    ROW090              -1
    COL158    ROW001             -75   ROW065             100
    COL158    ROW091              -1   ROW092              -1
    COL158    ROW093              -1
    COL159    ROW001             -75   ROW066             100
    COL159    ROW094              -1   ROW095              -1
    COL159    ROW096              -1
    MARK0001  'MARKER'                 'INTEND'
    COL160    ROW097               1   ROW001               1
RHS
    GRHS      ROW002               1   ROW003               1
    GRHS      ROW004               1   ROW005               1
    GRHS      ROW006               1   ROW007             200
    GRHS      ROW008             200   ROW009             200
    GRHS      ROW010               1   ROW011               1
    GRHS      ROW012               1   ROW013               1
    GRHS      ROW014               1   ROW015               1
    GRHS      ROW016               1   ROW017               1
    GRHS      ROW018               1   ROW019               1
    GRHS      ROW020               1   ROW021               1
    GRHS      ROW022               1   ROW023               1
    GRHS      ROW024               1   ROW025               1
    GRHS      ROW026               1   ROW027               1
    GRHS      ROW028               1   ROW029               1
    GRHS      ROW030               1   ROW031               1
    GRHS      ROW032               1   ROW033               1
    GRHS      ROW034               1   ROW035               1
    GRHS      ROW036               1   ROW037               1
    GRHS      ROW038               1   ROW039               1
    GRHS      ROW040               1   ROW062               1
    GRHS      ROW063               1   ROW064               1
    GRHS      ROW067               1   ROW068               1
    GRHS      ROW069               1   ROW070               1
    GRHS      ROW071               1   ROW072               1
    GRHS      ROW073               1   ROW074               1
    GRHS      ROW075               1   ROW076               1
    GRHS      ROW077               1   ROW078               1
    GRHS      ROW079               1   ROW080               1
    GRHS      ROW081               1   ROW082               1
    GRHS      ROW083               1   ROW084               1
    GRHS      ROW085               1   ROW086               1
    GRHS      ROW087               1   ROW088               1
    GRHS      ROW089               1   ROW090               1
    GRHS      ROW091               1   ROW092               1
    GRHS      ROW093               1   ROW094               1
    GRHS      ROW095               1   ROW096               1
BOUNDS
 UP GBND      COL001               1
 UP GBND      COL002               1
 UP GBND      COL003               1
 UP GBND      COL004               1
 UP GBND      COL005               1
 UP GBND      COL006               1
 UP GBND      COL007               1
 UP GBND      COL008               1
 UP GBND      COL009               1
 UP GBND      COL010               1
 UP GBND      COL011               1
 UP GBND      COL012               1
 UP GBND      COL013               1
 UP GBND      COL014               1
 UP GBND      COL015               1
 UP GBND      COL016               1
 UP GBND      COL017               1
 UP GBND      COL018               1
 UP GBND      COL019               1
 UP GBND      COL020               1
 UP GBND      COL021               1
 UP GBND      COL022               1
 UP GBND      COL023               1
 UP GBND      COL024               1
 UP GBND      COL025               1
 UP GBND      COL026               1
 UP GBND      COL027               1
 UP GBND      COL028               1
 UP GBND      COL029               1
 UP GBND      COL030               1
 UP GBND      COL031               1
 UP GBND      COL032               1
 UP GBND      COL033               1
 UP GBND      COL034               1
 UP GBND      COL035               1
 UP GBND      COL036               1
 UP GBND      COL037               1
 UP GBND      COL038               1
 UP GBND      COL039               1
 UP GBND      COL040               1
 UP GBND      COL041               1
 UP GBND      COL042               1
 UP GBND      COL043               1
 UP GBND      COL044               1
 UP GBND      COL045               1
 UP GBND      COL046               1
 UP GBND      COL047               1
 UP GBND      COL048               1
 UP GBND      COL049               1
 UP GBND      COL050               1
 UP GBND      COL051               1
 UP GBND      COL052               1
 UP GBND      COL053               1
 UP GBND      COL054               1
 UP GBND      COL055               1
 UP GBND      COL056               1
 UP GBND      COL057               1
 UP GBND      COL058               1
 UP GBND      COL059               1
 UP GBND      COL060               1
 UP GBND      COL061               1
 UP GBND      COL062               1
 UP GBND      COL063               1
 UP GBND      COL064               1
 UP GBND      COL065               1
 UP GBND      COL066               1
 UP GBND      COL067               1
 UP GBND      COL068               1
 UP GBND      COL069               1
 UP GBND      COL070               1
 UP GBND      COL071               1
 UP GBND      COL072               1
 UP GBND      COL073               1
 UP GBND      COL074               1
 UP GBND      COL075               1
 UP GBND      COL076               1
 UP GBND      COL077               1
 UP GBND      COL078               1
 UP GBND      COL079               1
 UP GBND      COL080               1
 UP GBND      COL081               1
 UP GBND      COL082               1
 UP GBND      COL083               1
 UP GBND      COL084               1
 UP GBND      COL085               1
 UP GBND      COL086               1
 UP GBND      COL087               1
 UP GBND      COL088               1
 UP GBND      COL089               1
 UP GBND      COL090               1
 UP GBND      COL091               1
 UP GBND      COL092               1
 UP GBND      COL093               1
 UP GBND      COL094               1
 UP GBND      COL095               1
 UP GBND      COL096               1
 UP GBND      COL097               1
 UP GBND      COL098               1
 UP GBND      COL099               1
 UP GBND      COL100               1
 UP GBND      COL101               1
 UP GBND      COL102               1
 UP GBND      COL103               1
 UP GBND      COL104               1
 UP GBND      COL105               1
 UP GBND      COL106               1
 UP GBND      COL107               1
 UP GBND      COL108               1
 UP GBND      COL109               1
 UP GBND      COL110               1
 UP GBND      COL111               1
 UP GBND      COL112               1
 UP GBND      COL113               1
 UP GBND      COL114               1
 UP GBND      COL115               1
 UP GBND      COL116               1
 UP GBND      COL117               1
 UP GBND      COL118               1
 UP GBND      COL119               1
 UP GBND      COL120               1
 UP GBND      COL121               1
 UP GBND      COL122               1
 UP GBND      COL123               1
 UP GBND      COL124               1
 UP GBND      COL125               1
 UP GBND      COL126               1
 UP GBND      COL127               1
 UP GBND      COL128               1
 UP GBND      COL129               1
 UP GBND      COL130               1
 UP GBND      COL131               1
 UP GBND      COL132               1
 UP GBND      COL133               1
 UP GBND      COL134               1
 UP GBND      COL135               1
 UP GBND      COL136               1
 UP GBND      COL137               1
 UP GBND      COL138               1
 UP GBND      COL139               1
 UP GBND      COL140               1
 UP GBND      COL141               1
 UP GBND      COL142               1
 UP GBND      COL143               1
 UP GBND      COL144               1
 UP GBND      COL145               1
 UP GBND      COL146               1
 UP GBND      COL147               1
 UP GBND      COL148               1
 UP GBND      COL149               1
 UP GBND      COL150               1
 UP GBND      COL151               1
 UP GBND      COL152               1
 UP GBND      COL153               1
 UP GBND      COL154               1
 UP GBND      COL155               1
 UP GBND      COL156               1
 UP GBND      COL157               1
 UP GBND      COL158               1
 UP GBND      COL159               1
 FR GBND      COL160  
ENDATA
